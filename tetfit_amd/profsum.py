"""Summaries of rocprofv3 CSV output (``-f csv``) for the kernels of this
package: per-kernel dispatch statistics from a ``--kernel-trace`` run and HBM
traffic per dispatch from ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs,
scaled by what the same counters read for ``bin/hbm_copy`` (a copy of known
size).

    python -m tetfit_amd.profsum trace  <trace_dir> [--min-us 5] [--match REGEX]
    python -m tetfit_amd.profsum traffic <pmc_fetch_dir> <pmc_write_dir> \
        <copy_fetch_dir> <copy_write_dir> [--min-kib 1000] [--match REGEX]
    python -m tetfit_amd.profsum requests <pmc_dir> [<pmc_dir> ...] [--match REGEX]

``requests`` reads runs that collected the TCC/EA request counters by size
(TCC_EA0_RDREQ_sum, _32B_sum, _64B_sum, _128B_sum; TCC_BUBBLE_sum,
TCC_EA0_WRREQ_sum, TCC_EA0_WRREQ_64B_sum) and reports each kernel's request
mix with the bytes it implies (32, 64 or 128 per read request, 64 or 32 per
write request) beside the byte count of the FETCH_SIZE expression, which
weighs 128-byte reads by TCC_BUBBLE.

Dispatches shorter than ``--min-us`` (trace) or below ``--min-kib`` of
fetched data (traffic) are CG passes in which the stop test fired: the kernel
returns after the test, so those dispatches do no matrix work and are listed
apart from the working ones.
"""
import argparse
import collections
import csv
import glob
import json
import os
import re
import sys


def _one(directory, suffix):
    files = sorted(glob.glob(os.path.join(directory, "*" + suffix)))
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    return files[0]


def _short(name):
    return name.split("(")[0].replace("void ", "")


def trace_stats(directory, min_us=5.0, match=None):
    """{kernel: dict(n, mean_us, n_work, mean_work_us, total_ms)} from a kernel_trace.csv."""
    d = collections.defaultdict(list)
    with open(_one(directory, "kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            k = _short(r["Kernel_Name"])
            if match and not re.search(match, k):
                continue
            d[k].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0)
    out = {}
    for k, v in d.items():
        w = [x for x in v if x >= min_us]
        out[k] = dict(n=len(v), mean_us=sum(v) / len(v), n_work=len(w),
                      mean_work_us=(sum(w) / len(w)) if w else None, total_ms=sum(v) / 1000.0)
    return out


def counter_means(directory, min_kib=0.0, match=None):
    d = collections.defaultdict(list)
    with open(_one(directory, "counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            k = _short(r["Kernel_Name"])
            if match and not re.search(match, k):
                continue
            d[k].append(float(r["Counter_Value"]))
    out = {}
    for k, v in d.items():
        w = [x for x in v if x >= min_kib]
        out[k] = dict(n=len(v), n_work=len(w), mean_kib=(sum(w) / len(w)) if w else None)
    return out


def counter_table(dirs, match=None, work_frac=0.05):
    """{kernel: {counter: mean per dispatch}} over the working dispatches of
    one or more --pmc runs (each run's counter_collection.csv). Within a run,
    a dispatch whose counters all stay below work_frac of the kernel's
    largest values is a CG pass whose stop test fired (it returns before
    the matrix work) and is left out; 'dispatches' counts those kept."""
    out = collections.defaultdict(dict)
    for directory in dirs:
        disp = collections.defaultdict(lambda: collections.defaultdict(dict))  # kernel -> dispatch -> counter
        with open(_one(directory, "counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                k = _short(r["Kernel_Name"])
                if match and not re.search(match, k):
                    continue
                disp[k][r["Dispatch_Id"]][r["Counter_Name"]] = float(r["Counter_Value"])
        for k, by in disp.items():
            names = sorted({c for v in by.values() for c in v})
            peak = {c: max(v.get(c, 0.0) for v in by.values()) for c in names}
            keep = [v for v in by.values() if any(v.get(c, 0.0) >= work_frac * peak[c] and peak[c] > 0 for c in names)]
            keep = keep or list(by.values())
            for c in names:
                out[k][c] = sum(v.get(c, 0.0) for v in keep) / len(keep)
            out[k]["dispatches"] = max(out[k].get("dispatches", 0), len(keep))
    return dict(out)


def request_bytes(c):
    """Bytes per dispatch from the request counters of one kernel (means)."""
    g = lambda n: c.get(n, 0.0)
    rd, r32, r64, r128 = g("TCC_EA0_RDREQ_sum"), g("TCC_EA0_RDREQ_32B_sum"), g("TCC_EA0_RDREQ_64B_sum"), g(
        "TCC_EA0_RDREQ_128B_sum")
    wr, w64 = g("TCC_EA0_WRREQ_sum"), g("TCC_EA0_WRREQ_64B_sum")
    bub = g("TCC_BUBBLE_sum")
    out = dict(dispatches=int(g("dispatches")), rdreq=rd, rd32=r32, rd64=r64, rd128=r128, bubble=bub, wrreq=wr, wr64=w64,
               read_bytes=32 * r32 + 64 * r64 + 128 * r128, write_bytes=64 * w64 + 32 * (wr - w64))
    # the FETCH_SIZE expression: 128 B per TCC_BUBBLE, 32 B per 32-byte request, 64 B for the rest
    out["fetch_size_bytes"] = 128 * bub + 64 * (rd - bub - r32) + 32 * r32
    return out


def traffic(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, min_kib=1000.0, match=None):
    """Bytes per dispatch read from and written to memory beyond L2, scaled
    so that the copy kernel's counters give its known byte counts."""
    cf = counter_means(copy_fetch_dir, match="copy_kernel")["copy_kernel"]["mean_kib"]
    cw = counter_means(copy_write_dir, match="copy_kernel")["copy_kernel"]["mean_kib"]
    copy_bytes = 1 << 30  # hbm_copy's default buffer: 1 GiB read and 1 GiB written per launch
    f_scale = copy_bytes / (cf * 1024.0)
    w_scale = copy_bytes / (cw * 1024.0)
    fe = counter_means(fetch_dir, min_kib, match)
    wr = counter_means(write_dir, min_kib / 16.0, match)
    out = {}
    for k in sorted(set(fe) & set(wr)):
        if fe[k]["mean_kib"] is None or wr[k]["mean_kib"] is None:
            continue
        rb = fe[k]["mean_kib"] * 1024.0 * f_scale
        wb = wr[k]["mean_kib"] * 1024.0 * w_scale
        out[k] = dict(read_bytes=rb, write_bytes=wb, traffic=rb + wb, dispatches=fe[k]["n_work"])
    return out, dict(fetch_scale=f_scale, write_scale=w_scale, copy_fetch_kib=cf, copy_write_kib=cw)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--min-us", type=float, default=5.0)
    t.add_argument("--match", default=None)
    t.add_argument("--top", type=int, default=25)
    p = sub.add_parser("traffic")
    p.add_argument("fetch")
    p.add_argument("write")
    p.add_argument("copy_fetch")
    p.add_argument("copy_write")
    p.add_argument("--min-kib", type=float, default=1000.0)
    p.add_argument("--match", default=None)
    p.add_argument("--json", action="store_true")
    q = sub.add_parser("requests")
    q.add_argument("dirs", nargs="+")
    q.add_argument("--match", default=None)
    q.add_argument("--json", action="store_true")
    a = ap.parse_args(argv)
    if a.cmd == "requests":
        tab = {k: request_bytes(c) for k, c in counter_table(a.dirs, a.match).items()}
        if a.json:
            print(json.dumps(tab))
            return
        print("| kernel | dispatches | read req (M) | 32 B | 64 B | 128 B | TCC_BUBBLE (M) | read MB (by size) | "
              "FETCH_SIZE expr. MB | ratio | write req (M) | 64 B | write MB |")
        print("|---|---|---|---|---|---|---|---|---|---|---|---|---|")
        for k, v in sorted(tab.items()):
            rd = max(v["rdreq"], 1.0)
            print(f"| `{k}` | {v['dispatches']} | {v['rdreq'] / 1e6:.3f} | {v['rd32'] / rd:.1%} | {v['rd64'] / rd:.1%} | "
                  f"{v['rd128'] / rd:.1%} | {v['bubble'] / 1e6:.3f} | {v['read_bytes'] / 1e6:.2f} | "
                  f"{v['fetch_size_bytes'] / 1e6:.2f} | "
                  f"{v['read_bytes'] / max(v['fetch_size_bytes'], 1.0):.3f} | {v['wrreq'] / 1e6:.3f} | "
                  f"{v['wr64'] / max(v['wrreq'], 1.0):.1%} | {v['write_bytes'] / 1e6:.2f} |")
        return
    if a.cmd == "trace":
        st = trace_stats(a.dir, a.min_us, a.match)
        print("| kernel | dispatches | mean µs (all) | working dispatches (≥ %g µs) | mean µs (working) | total ms |" % a.min_us)
        print("|---|---|---|---|---|---|")
        for k, v in sorted(st.items(), key=lambda kv: -kv[1]["total_ms"])[: a.top]:
            mw = "%.2f" % v["mean_work_us"] if v["mean_work_us"] is not None else "—"
            print(f"| `{k}` | {v['n']} | {v['mean_us']:.2f} | {v['n_work']} | {mw} | {v['total_ms']:.3f} |")
    else:
        tr, cal = traffic(a.fetch, a.write, a.copy_fetch, a.copy_write, a.min_kib, a.match)
        if a.json:
            print(json.dumps(dict(kernels=tr, calibration=cal)))
            return
        print(f"calibration: FETCH_SIZE x {cal['fetch_scale']:.4f}, WRITE_SIZE x {cal['write_scale']:.4f} "
              f"(copy of 1 GiB read 1 GiB written read {cal['copy_fetch_kib']:.0f} / {cal['copy_write_kib']:.0f} KiB)")
        print("| kernel | dispatches | read MB | written MB | traffic MB |")
        print("|---|---|---|---|---|")
        for k, v in tr.items():
            print(f"| `{k}` | {v['dispatches']} | {v['read_bytes'] / 1e6:.2f} | {v['write_bytes'] / 1e6:.2f} | "
                  f"{v['traffic'] / 1e6:.2f} |")


if __name__ == "__main__":
    sys.exit(main())
