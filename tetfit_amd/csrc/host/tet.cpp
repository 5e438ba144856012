// tet.cpp - netgen tetrahedral mesh reader and the DGA operator producers.
// Restates ReadUnstructuredMesh (src/DGA/Discretization.hpp:5868-6926):
// node-sorted tets sorted by node tuple, faces and edges numbered by
// sort+unique of node tuples, signed incidences (fte +1,-1,+1), boundary dual
// edge vectors and Ctb, pec / source marking and edge classes; then
// ConstructCodecasaMaterialMatrices (D:8694-9513, method dga) and
// ConstructGFDTDMaterialMatrices (D:8132-8692, method gfdtd), FindProbe
// (D:3645-3747) and the Whitney / face probe weights (D:3768-3897).
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <parallel/algorithm>
#include <sstream>
#include <tuple>
#include <type_traits>

#include "build.hpp"

namespace tfh {

namespace {

std::vector<char> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw Error("cannot open mesh file " + path);
    f.seekg(0, std::ios::end);
    size_t n = (size_t)f.tellg();
    f.seekg(0);
    std::vector<char> buf(n + 1);
    f.read(buf.data(), (std::streamsize)n);
    buf[n] = 0;
    return buf;
}

struct Reader {
    std::vector<char> buf;
    const char* p = nullptr;
    const char* end = nullptr;
    explicit Reader(std::vector<char> text) : buf(std::move(text)) {
        if (buf.empty() || buf.back() != 0) buf.push_back(0);
        p = buf.data();
        end = p + buf.size() - 1;
    }
    double num() {
        char* e;
        double v = std::strtod(p, &e);
        if (e == p) throw Error("malformed netgen mesh file");
        p = e;
        return v;
    }
    uint64_t uint() {
        char* e;
        unsigned long long v = std::strtoull(p, &e, 10);
        if (e == p) throw Error("malformed netgen mesh file");
        p = e;
        return v;
    }
};

// ConvertFromGMSH (D:5692-5866): a gmsh 2.2 ASCII mesh turned into the
// netgen text the reader takes, line for line as the reference writes its
// temporary .mesh file: after the 4 header lines the node count and the
// nodes (id x y z, written back as "x y z" at the stream's default 6
// significant digits), 2 more lines, the element count and the elements
// (number, type, tag count, tags, nodes; read_element_line, Utilities.hpp:
// 159-223 takes the first two numbers after the type as tags, the first of
// them the label): tetrahedra (type 4) as "label n0 n1 n2 n3" with the
// nodes sorted and the tetrahedra sorted by node tuple, triangles (type 2)
// as "label n0 n1 n2" in file order.
std::vector<char> convert_gmsh(const std::string& path) {
    std::vector<char> in = read_file(path);
    const char* p = in.data();
    auto next_line = [&]() {
        while (*p && *p != '\n') ++p;
        if (*p) ++p;
    };
    auto num = [&](auto& v) {
        char* e;
        if constexpr (std::is_floating_point<std::decay_t<decltype(v)>>::value) v = std::strtod(p, &e);
        else v = (std::decay_t<decltype(v)>)std::strtoll(p, &e, 10);
        if (e == p) throw Error("malformed gmsh mesh file " + path);
        p = e;
    };
    for (int i = 0; i < 4; ++i) next_line();
    uint32_t npts = 0;
    num(npts);
    std::string out = std::to_string(npts) + "\n";
    next_line();
    char buf[128];
    for (uint32_t i = 0; i < npts; ++i) {
        double id, x, y, z;
        num(id); num(x); num(y); num(z);
        std::snprintf(buf, sizeof buf, "%g %g %g\n", x, y, z);
        out += buf;
        next_line();
    }
    for (int i = 0; i < 2; ++i) next_line();
    uint32_t nel = 0;
    num(nel);
    struct TetL { uint32_t n[4]; uint32_t lab; };
    std::vector<TetL> tets;
    std::vector<std::array<uint32_t, 4>> tris;  // label, n0, n1, n2
    for (uint32_t i = 0; i < nel; ++i) {
        next_line();
        int64_t number, type;
        num(number); num(type);
        if (type == 4 || type == 2) {
            int64_t ntags, tag1, tag2;
            num(ntags); num(tag1); num(tag2);
            const int nn = type == 4 ? 4 : 3;
            int64_t v[4];
            for (int k = 0; k < nn; ++k) num(v[k]);
            if (type == 4) {
                TetL t;
                for (int k = 0; k < 4; ++k) t.n[k] = (uint32_t)v[k];
                std::sort(t.n, t.n + 4);
                t.lab = (uint32_t)tag1;
                tets.push_back(t);
            } else {
                tris.push_back({(uint32_t)tag1, (uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2]});
            }
        }
    }
    std::stable_sort(tets.begin(), tets.end(), [](const TetL& a, const TetL& c) {
        return std::lexicographical_compare(a.n, a.n + 4, c.n, c.n + 4);
    });
    out += std::to_string(tets.size()) + "\n";
    for (auto& t : tets) {
        std::snprintf(buf, sizeof buf, "%u %u %u %u %u\n", t.lab, t.n[0], t.n[1], t.n[2], t.n[3]);
        out += buf;
    }
    out += std::to_string(tris.size()) + "\n";
    for (auto& t : tris) {
        std::snprintf(buf, sizeof buf, "%u %u %u %u\n", t[0], t[1], t[2], t[3]);
        out += buf;
    }
    return std::vector<char>(out.begin(), out.end());
}

struct Tri { uint32_t a, b, c, slot; };
struct Ed { uint32_t a, b, slot; };

// Cholesky solve of an SPD matrix (n x n, row-major) for its inverse.
// Inverse of a general n x n matrix (row-major) by Gauss-Jordan elimination
// with partial pivoting; false when singular.
bool general_inverse(const std::vector<double>& A0, int n, std::vector<double>& inv) {
    std::vector<double> A(A0);
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)piv * n + c])) piv = r;
        if (!(A[(size_t)piv * n + c] != 0)) return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(A[(size_t)c * n + k], A[(size_t)piv * n + k]);
                std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
            }
        const double d = A[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) { A[(size_t)c * n + k] /= d; inv[(size_t)c * n + k] /= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = A[(size_t)r * n + c];
            if (f == 0) continue;
            for (int k = 0; k < n; ++k) {
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
                inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
            }
        }
    }
    return true;
}

bool spd_inverse(const std::vector<double>& A, int n, std::vector<double>& inv) {
    std::vector<double> L(n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        double s = A[j * n + j];
        for (int k = 0; k < j; ++k) s -= L[j * n + k] * L[j * n + k];
        if (!(s > 0)) return false;
        double d = std::sqrt(s);
        L[j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double t = A[i * n + j];
            for (int k = 0; k < j; ++k) t -= L[i * n + k] * L[j * n + k];
            L[i * n + j] = t / d;
        }
    }
    inv.assign(n * n, 0.0);
    std::vector<double> y(n), x(n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {
            double t = i == c ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) t -= L[i * n + k] * y[k];
            y[i] = t / L[i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double t = y[i];
            for (int k = i + 1; k < n; ++k) t -= L[k * n + i] * x[k];
            x[i] = t / L[i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[i * n + c] = x[i];
    }
    return true;
}

inline double& csr_at(CsrStore& m, int32_t r, int32_t c) {
    auto b = m.col.begin() + m.rowptr[r], e = m.col.begin() + m.rowptr[r + 1];
    auto it = std::lower_bound(b, e, c);
    if (it == e || *it != c) throw Error("internal: sparse pattern miss");
    return m.val[it - m.col.begin()];
}

// Locality ranks of the edges and faces (tf_dga_ops.edge_rank, face_rank).
// The vertices are snapped to a lattice of spacing h per axis (h from the
// median largest component of the edge vectors, adjusted so the bounding box
// holds a whole number of steps that the vertices sit closest to; on a
// jittered Kuhn mesh this recovers the lattice exactly). An edge or face
// gets a base lattice point (the first of its snapped vertices in (z, y, x)
// order) and a shape class (the other vertices' lattice offsets from it),
// and the entities are sorted by (z, band of J y-lines, class, y, x):
// entities of one class along one x-line are consecutive, so a 64-row slice
// of same-class rows reaches each neighbour class through consecutive
// positions. J bounds a (z, band) block to about 8,192 entities, so a row's
// neighbours in the planes z-1, z, z+1 stay in three narrow position ranges.
struct Lattice {
    double lo[3] = {0, 0, 0}, h[3] = {1, 1, 1};
    int64_t cnt[3] = {1, 1, 1};  // lattice points per axis
    bool ok = false;
    void snap(const Vec3& p, int64_t* q) const {
        const double c[3] = {p.x, p.y, p.z};
        for (int a = 0; a < 3; ++a)
            q[a] = std::min<int64_t>(cnt[a] - 1, std::max<int64_t>(0, std::llround((c[a] - lo[a]) / h[a])));
    }
};

Lattice mesh_lattice(const std::vector<Vec3>& pts, const std::vector<uint32_t>& en, size_t E) {
    Lattice L;
    if (E == 0 || pts.empty()) return L;
    double hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int a = 0; a < 3; ++a) L.lo[a] = HUGE_VAL;
    for (const Vec3& p : pts) {
        const double c[3] = {p.x, p.y, p.z};
        for (int a = 0; a < 3; ++a) { L.lo[a] = std::min(L.lo[a], c[a]); hi[a] = std::max(hi[a], c[a]); }
    }
    std::vector<double> big;
    const size_t stride = std::max<size_t>(1, E / 200000);
    for (size_t e = 0; e < E; e += stride) {
        const Vec3 d = pts[en[2 * e + 1]] - pts[en[2 * e]];
        big.push_back(std::max(std::fabs(d.x), std::max(std::fabs(d.y), std::fabs(d.z))));
    }
    std::nth_element(big.begin(), big.begin() + big.size() / 2, big.end());
    const double hmed = big[big.size() / 2];
    if (!(hmed > 0)) return L;
    // steps per axis: of the counts within [0.75, 1.35] x extent/hmed (the
    // median is biased high by the diagonal edges), the one whose lattice the
    // vertices sit closest to (sum of squared distances over a sample)
    const size_t vstride = std::max<size_t>(1, pts.size() / 20000);
    for (int a = 0; a < 3; ++a) {
        const double ext = hi[a] - L.lo[a];
        const int64_t c0 = std::max<int64_t>(1, std::llround(ext / hmed));
        int64_t best = c0;
        double best_r = HUGE_VAL;
        for (int64_t c = std::max<int64_t>(1, (int64_t)(0.75 * c0)); ext > 0 && c <= (int64_t)(1.35 * c0) + 1; ++c) {
            const double s = (double)c / ext;
            double r = 0;
            for (size_t i = 0; i < pts.size(); i += vstride) {
                const double u = ((a == 0 ? pts[i].x : a == 1 ? pts[i].y : pts[i].z) - L.lo[a]) * s;
                r += (u - std::nearbyint(u)) * (u - std::nearbyint(u));
            }
            if (r < best_r) { best_r = r; best = c; }
        }
        L.h[a] = ext > 0 ? ext / (double)best : 1.0;
        L.cnt[a] = best + 1;
    }
    L.ok = true;
    return L;
}

// Rank of n entities of `per` (2: edges, 3: faces) vertices each; empty when
// the lattice is unusable or the keys would not fit 63 bits.
std::vector<int32_t> locality_rank(const Lattice& L, const std::vector<Vec3>& pts, const std::vector<uint32_t>& nodes,
                                   int per, size_t n) {
    std::vector<int32_t> rank;
    if (!L.ok || n == 0) return rank;
    auto enc = [](const int64_t* o) {
        int c = 0, m = 1;
        for (int a = 0; a < 3; ++a, m *= 5) c += (int)(std::min<int64_t>(2, std::max<int64_t>(-2, o[a])) + 2) * m;
        return c;
    };
    std::vector<int32_t> code(n);
    std::vector<int64_t> base(3 * n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        std::array<std::array<int64_t, 3>, 3> q{};
        for (int k = 0; k < per; ++k) L.snap(pts[nodes[per * i + k]], q[k].data());
        // (z, y, x) order of the snapped vertices
        std::sort(q.begin(), q.begin() + per, [](const std::array<int64_t, 3>& u, const std::array<int64_t, 3>& v) {
            return std::make_tuple(u[2], u[1], u[0]) < std::make_tuple(v[2], v[1], v[0]);
        });
        int c = 0;
        for (int k = 1; k < per; ++k) {
            const int64_t o[3] = {q[k][0] - q[0][0], q[k][1] - q[0][1], q[k][2] - q[0][2]};
            c = c * 125 + enc(o);
        }
        code[i] = c;
        for (int a = 0; a < 3; ++a) base[3 * i + a] = q[0][a];
    }
    // classes numbered densely in code order
    std::vector<int32_t> codes(code);
    std::sort(codes.begin(), codes.end());
    codes.erase(std::unique(codes.begin(), codes.end()), codes.end());
    const int64_t ncls = (int64_t)codes.size();
    const int64_t J = std::max<int64_t>(1, 8192 / std::max<int64_t>(1, ncls * L.cnt[0]));
    const int64_t nband = (L.cnt[1] + J - 1) / J;
    const long double span = (long double)L.cnt[2] * nband * ncls * L.cnt[1] * L.cnt[0];
    if (span >= 4.6e18L) return rank;
    std::vector<std::pair<int64_t, int32_t>> key(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const int64_t x = base[3 * i], y = base[3 * i + 1], z = base[3 * i + 2];
        const int64_t cls = std::lower_bound(codes.begin(), codes.end(), code[i]) - codes.begin();
        key[i] = {(((z * nband + y / J) * ncls + cls) * L.cnt[1] + y) * L.cnt[0] + x, (int32_t)i};
    }
    // (key, index) pairs are distinct, so any sort gives the same order
    __gnu_parallel::sort(key.begin(), key.end());
    rank.resize(n);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)n; ++r) rank[key[r].second] = (int32_t)r;
    return rank;
}

}  // namespace

void build_tet(const Model& m, const SimulationDef& sim, Build& b) {
    auto mit = m.meshes.find(sim.mesh);
    if (mit == m.meshes.end()) throw Error("simulation refers to an undefined mesh");
    const MeshDef& msh = mit->second;

    for (auto& kv : m.bcs)
        if (kv.second.type == "pml") throw Error("PML boundary conditions are not supported by this build");
    std::map<uint32_t, Material> materials = m.materials;
    std::map<uint32_t, BoundaryCondition> bcs = m.bcs;
    std::map<uint32_t, SourceDef> sources = m.sources;
    const bool is_dga = sim.method == "dga";
    const bool is_fem = sim.method == "fem";
    const double scale = msh.scale;

    PhaseClock clk;
    Reader rd(msh.mesher == "gmsh" ? convert_gmsh(msh.file) : read_file(msh.file));
    // ---- points
    const uint64_t np = rd.uint();
    std::vector<Vec3> pts(np);
    for (uint64_t i = 0; i < np; ++i) {
        double x = rd.num(), y = rd.num(), z = rd.num();
        pts[i] = Vec3(x, y, z) * scale;
    }
    clk.mark("read points");
    // ---- tetrahedra (node-sorted, label 7 dropped, sorted by node tuple)
    const uint64_t nt_file = rd.uint();
    struct TetIn { uint32_t n[4]; uint32_t mat; };
    std::vector<TetIn> tin;
    tin.reserve(nt_file);
    for (uint64_t i = 0; i < nt_file; ++i) {
        TetIn t;
        t.mat = (uint32_t)rd.uint();
        for (int k = 0; k < 4; ++k) {
            uint64_t v = rd.uint();
            if (v < 1 || v > np) throw Error("netgen mesh: tetrahedron node out of range");
            t.n[k] = (uint32_t)(v - 1);
        }
        std::sort(t.n, t.n + 4);
        if (t.mat != 7) tin.push_back(t);
    }
    std::sort(tin.begin(), tin.end(), [](const TetIn& a, const TetIn& c) {
        return std::lexicographical_compare(a.n, a.n + 4, c.n, c.n + 4);
    });
    const uint32_t T = (uint32_t)tin.size();
    if (T == 0) throw Error("netgen mesh has no tetrahedra");
    b.vols.resize((size_t)T * 4);
    b.vol_material.resize(T);
    b.cell_volumes.resize(T);
    std::vector<Mat3> dvp_eps(np), dvp_sig(np);
    std::vector<uint8_t> cls_node(np, 0);
    std::vector<uint32_t> av_cnt(np + 1, 0);
    std::vector<Tri> tri((size_t)T * 4);
    for (uint32_t t = 0; t < T; ++t) {
        const uint32_t* n = tin[t].n;
        for (int k = 0; k < 4; ++k) { b.vols[(size_t)t * 4 + k] = (int32_t)n[k]; av_cnt[n[k] + 1]++; }
        b.vol_material[t] = (int32_t)tin[t].mat;
        tri[t] = {n[0], n[1], n[2], t};
        tri[t + T] = {n[0], n[1], n[3], t + T};
        tri[t + 2 * (size_t)T] = {n[0], n[2], n[3], t + 2 * T};
        tri[t + 3 * (size_t)T] = {n[1], n[2], n[3], t + 3 * T};
        Vec3 v1 = pts[n[1]] - pts[n[0]], v2 = pts[n[2]] - pts[n[0]], v3 = pts[n[3]] - pts[n[0]];
        b.cell_volumes[t] = dot(v1, cross(v2, v3)) / 6.0;
        const Material& mt = materials[tin[t].mat];
        if (mt.eps_random) throw Error("epsilon_random is supported on cartesian meshes only");
        for (int k = 0; k < 4; ++k) {
            uint32_t p = n[k];
            if (!dvp_eps[p].is_zero()) {
                if (mt.sigma != dvp_sig[p]) cls_node[p] |= 4;
                if (mt.epsilon != dvp_eps[p]) cls_node[p] |= 2;
                if (!dvp_sig[p].is_zero()) cls_node[p] |= 1;
            } else {
                dvp_eps[p] = mt.epsilon;
                dvp_sig[p] = mt.sigma;
                if (!dvp_sig[p].is_zero()) cls_node[p] |= 1;
            }
        }
    }
    std::vector<TetIn>().swap(tin);
    // associated_volumes as CSR (tets in increasing order per node)
    for (uint64_t i = 0; i < np; ++i) av_cnt[i + 1] += av_cnt[i];
    std::vector<uint32_t> av(av_cnt[np]);
    {
        std::vector<uint32_t> pos(av_cnt.begin(), av_cnt.end() - 1);
        for (uint32_t t = 0; t < T; ++t)
            for (int k = 0; k < 4; ++k) av[pos[b.vols[(size_t)t * 4 + k]]++] = t;
    }

    clk.mark("tets");
    // ---- faces: sort + unique (unique() D:10173-10219)
    std::vector<uint32_t> flab((size_t)T * 4);
    // (copies of one face are equal under the comparator and get one label,
    // so the order among them, which a parallel sort leaves open, is moot)
    __gnu_parallel::sort(tri.begin(), tri.end(), [](const Tri& x, const Tri& y) {
        if (x.a != y.a) return x.a < y.a;
        if (x.b != y.b) return x.b < y.b;
        return x.c < y.c;
    });
    std::vector<uint32_t> fn;
    {
        uint32_t it = 0;
        flab[tri[0].slot] = 0;
        fn.push_back(tri[0].a); fn.push_back(tri[0].b); fn.push_back(tri[0].c);
        for (size_t k = 1; k < tri.size(); ++k) {
            if (!(tri[k].a == tri[k - 1].a && tri[k].b == tri[k - 1].b && tri[k].c == tri[k - 1].c)) {
                ++it;
                fn.push_back(tri[k].a); fn.push_back(tri[k].b); fn.push_back(tri[k].c);
            }
            flab[tri[k].slot] = it;
        }
    }
    std::vector<Tri>().swap(tri);
    const uint32_t F = (uint32_t)(fn.size() / 3);
    // ftv (up to 2 tets per face, with signs) and vtf (4 faces per tet)
    std::vector<int32_t> ftv_t((size_t)F * 2, -1);
    std::vector<int8_t> ftv_s((size_t)F * 2, 0);
    std::vector<uint8_t> ftv_n(F, 0);
    b.vtf.resize((size_t)T * 4);
    std::vector<int8_t> vtf_sgn((size_t)T * 4);
    for (uint32_t k = 0; k < T; ++k) {
        const uint32_t l[4] = {flab[k], flab[k + T], flab[k + 2 * (size_t)T], flab[k + 3 * (size_t)T]};
        const int8_t s[4] = {-1, 1, -1, 1};  // vol_signs is always +1 (D:6186)
        for (int q = 0; q < 4; ++q) {
            uint8_t& n = ftv_n[l[q]];
            if (n >= 2) throw Error("non-manifold tetrahedral mesh: a face shared by more than 2 tets");
            ftv_t[(size_t)l[q] * 2 + n] = (int32_t)k;
            ftv_s[(size_t)l[q] * 2 + n] = s[q];
            ++n;
        }
        // vtf_list[k] = {f4, f3, f2, f1} (D:6263-6266)
        b.vtf[(size_t)k * 4 + 0] = (int32_t)l[3]; vtf_sgn[(size_t)k * 4 + 0] = 1;
        b.vtf[(size_t)k * 4 + 1] = (int32_t)l[2]; vtf_sgn[(size_t)k * 4 + 1] = -1;
        b.vtf[(size_t)k * 4 + 2] = (int32_t)l[1]; vtf_sgn[(size_t)k * 4 + 2] = 1;
        b.vtf[(size_t)k * 4 + 3] = (int32_t)l[0]; vtf_sgn[(size_t)k * 4 + 3] = -1;
    }
    std::vector<uint32_t>().swap(flab);

    clk.mark("faces");
    // ---- edges: sort + unique with etn / nte / dual_star_offsets (D:10221-10299)
    std::vector<Ed> ed((size_t)F * 3);
    for (uint32_t f = 0; f < F; ++f) {
        uint32_t p0 = fn[(size_t)f * 3], p1 = fn[(size_t)f * 3 + 1], p2 = fn[(size_t)f * 3 + 2];
        ed[f] = {p0, p1, f};
        ed[f + F] = {p0, p2, f + F};
        ed[f + 2 * (size_t)F] = {p1, p2, f + 2 * F};
    }
    __gnu_parallel::sort(ed.begin(), ed.end(), [](const Ed& x, const Ed& y) {
        return x.a != y.a ? x.a < y.a : x.b < y.b;
    });
    std::vector<uint32_t> elab((size_t)F * 3);
    std::vector<uint32_t> en;
    {
        uint32_t it = 0;
        elab[ed[0].slot] = 0;
        en.push_back(ed[0].a); en.push_back(ed[0].b);
        for (size_t k = 1; k < ed.size(); ++k) {
            if (!(ed[k].a == ed[k - 1].a && ed[k].b == ed[k - 1].b)) {
                ++it;
                en.push_back(ed[k].a); en.push_back(ed[k].b);
            }
            elab[ed[k].slot] = it;
        }
    }
    std::vector<Ed>().swap(ed);
    const uint32_t E = (uint32_t)(en.size() / 2);
    // nte (edges around a node, increasing edge id) as CSR
    std::vector<uint32_t> nte_ptr(np + 1, 0);
    for (uint32_t e = 0; e < E; ++e) { nte_ptr[en[2 * (size_t)e] + 1]++; nte_ptr[en[2 * (size_t)e + 1] + 1]++; }
    for (uint64_t i = 0; i < np; ++i) nte_ptr[i + 1] += nte_ptr[i];
    std::vector<uint32_t> nte(nte_ptr[np]);
    {
        std::vector<uint32_t> pos(nte_ptr.begin(), nte_ptr.end() - 1);
        for (uint32_t e = 0; e < E; ++e) {
            nte[pos[en[2 * (size_t)e]]++] = e;
            nte[pos[en[2 * (size_t)e + 1]]++] = e;
        }
    }
    // fte: 3 signed edges per face, +1,-1,+1 (D:6345-6352)
    std::vector<int32_t> fte((size_t)F * 3);
    for (uint32_t f = 0; f < F; ++f) {
        fte[(size_t)f * 3] = (int32_t)elab[f];
        fte[(size_t)f * 3 + 1] = (int32_t)elab[f + F];
        fte[(size_t)f * 3 + 2] = (int32_t)elab[f + 2 * (size_t)F];
    }
    std::vector<uint32_t>().swap(elab);

    auto P = [&](uint32_t i) -> const Vec3& { return pts[i]; };
    std::vector<Vec3> face_bars(F), edge_bars(E);
    for (uint32_t f = 0; f < F; ++f) {
        Vec3 bc = P(fn[(size_t)f * 3]) + P(fn[(size_t)f * 3 + 1]) + P(fn[(size_t)f * 3 + 2]);
        face_bars[f] = bc / 3.0;
    }
    for (uint32_t e = 0; e < E; ++e) edge_bars[e] = 0.5 * (P(en[2 * (size_t)e]) + P(en[2 * (size_t)e + 1]));
    auto vol_bar = [&](uint32_t t) {
        const int32_t* n = &b.vols[(size_t)t * 4];
        return (P(n[0]) + P(n[1]) + P(n[2]) + P(n[3])) / 4.0;
    };

    clk.mark("edges");
    // ---- C, boundary faces, bnd edges and Ctb, face classes (D:6320-6546)
    b.boundary_face.assign(F, 0);
    std::vector<int32_t> bnd_edges(E, -1);
    std::vector<Vec3> bnd_dual;
    std::vector<uint8_t> face_class(F, 1);
    for (uint32_t k = 0; k < F; ++k) {
        const int nv = ftv_n[k];
        bool recombine = true;
        if (nv == 2) {
            const Material& m1 = materials[b.vol_material[ftv_t[(size_t)k * 2]]];
            const Material& m2 = materials[b.vol_material[ftv_t[(size_t)k * 2 + 1]]];
            if (!m1.chi.is_zero() || !m2.chi.is_zero()) {
                if (m1.chi != m2.chi) recombine = false;
                else if (m1.mu != m2.mu) recombine = false;
            }
        } else if (nv == 1) {
            b.boundary_face[k] = 1;
            for (int q = 0; q < 3; ++q) {
                int32_t e = fte[(size_t)k * 3 + q];
                if (bnd_edges[e] < 0) {
                    bnd_edges[e] = (int32_t)bnd_dual.size();
                    bnd_dual.push_back(face_bars[k]);
                } else {
                    Vec3& v = bnd_dual[bnd_edges[e]];
                    v = face_bars[k] - v;
                    Vec3 edgvec = P(en[2 * (size_t)e + 1]) - P(en[2 * (size_t)e]);
                    Vec3 a = P(fn[(size_t)k * 3]), c1 = P(fn[(size_t)k * 3 + 1]), c2 = P(fn[(size_t)k * 3 + 2]);
                    Vec3 facvec = cross(c1 - a, c2 - a);
                    Vec3 volvec = vol_bar((uint32_t)ftv_t[(size_t)k * 2]) - face_bars[k];
                    if (dot(facvec, volvec) >= 0) facvec = -facvec;
                    if (dot(cross(edgvec, facvec), v) > 0) v = -v;
                    v = -v;
                }
            }
        } else {
            throw Error("Conductor boundary cannot be on mesh boundary!");
        }
        const Material& m0 = materials[b.vol_material[ftv_t[(size_t)k * 2]]];
        if (recombine && m0.chi.is_zero()) face_class[k] = 1;
        else if (recombine) face_class[k] = 2;
        else face_class[k] = 3;
    }
    b.n_bnd = (int32_t)bnd_dual.size();

    // ---- boundary triangles: labels, pec, sources (D:6599-6740)
    std::vector<uint32_t> edge_bcs(E, 0);
    std::vector<std::vector<uint32_t>> edge_src(E);
    std::vector<uint8_t> is_pec(E, 0);
    auto bc_type = [&](uint32_t label) -> const std::string& { return bcs[label].type; };
    const uint64_t ntri = rd.uint();
    for (uint64_t i = 0; i < ntri; ++i) {
        uint32_t bid = (uint32_t)rd.uint();
        uint32_t v[3];
        for (int q = 0; q < 3; ++q) v[q] = (uint32_t)(rd.uint() - 1);
        std::sort(v, v + 3);
        // lower_bound in the sorted face list
        uint32_t lo = 0, hi = F;
        while (lo < hi) {
            uint32_t mid = (lo + hi) / 2;
            const uint32_t* c = &fn[(size_t)mid * 3];
            bool less = c[0] != v[0] ? c[0] < v[0] : (c[1] != v[1] ? c[1] < v[1] : c[2] < v[2]);
            if (less) lo = mid + 1; else hi = mid;
        }
        if (lo >= F) continue;
        const uint32_t* c = &fn[(size_t)lo * 3];
        if (!(c[0] == v[0] && c[1] == v[1] && c[2] == v[2])) continue;
        const uint32_t fl = lo;
        b.boundary_face[fl] = (int32_t)bid;
        const uint32_t which_bc = bid < m.bnd_to_bc.size() ? m.bnd_to_bc[bid] : 0;
        if (which_bc != 0) {
            if (bc_type(which_bc) == "pec") {
                for (int q = 0; q < 3; ++q) {
                    int32_t e = fte[(size_t)fl * 3 + q];
                    edge_bcs[e] = which_bc;
                    edge_src[e].clear();
                    is_pec[e] = 1;
                    if (bnd_edges[e] >= 0) {
                        bnd_dual[bnd_edges[e]] = Vec3(0, 0, 0);
                        bnd_edges[e] = -1;
                    }
                }
            }
        } else {
            std::vector<int32_t> edgs;
            for (int q = 0; q < 3; ++q) {
                int32_t e = fte[(size_t)fl * 3 + q];
                if (!is_pec[e]) edgs.push_back(e);
            }
            for (uint32_t sl : sim.sources) {
                const SourceDef& src = sources[sl];
                if (std::find(src.surface_label.begin(), src.surface_label.end(), bid) == src.surface_label.end()) continue;
                if (src.st == "b") continue;  // face sources are not used by the time loops
                for (int32_t e : edgs)
                    if (std::find(edge_src[e].begin(), edge_src[e].end(), sl) == edge_src[e].end()) edge_src[e].push_back(sl);
            }
        }
    }

    clk.mark("boundary, sources");
    // ---- edge classes (D:6770-6905)
    b.edge_class.assign(E, 0);
    b.h_row.assign(E, -1);
    b.fixed_value.assign(E, 0.0);
    int32_t H_size = 0;
    // fractional indices of the lossy dga unknowns (D:6806-6893): p_index of
    // class-2 edges, boundary_index of class-4 edges, and the offset of each
    // fractured node star's block of U_c (dual_is_fractured), all handed out
    // in edge order
    std::vector<int32_t> p_index(E, -1), bnd_index(E, -1), star_off(np, -1);
    int32_t P_size = 0, Q_size = 0, B_size = 0;
    auto fracture_star = [&](uint32_t node) {
        if (star_off[node] < 0) {
            star_off[node] = Q_size;
            Q_size += (int32_t)(nte_ptr[node + 1] - nte_ptr[node]);
        }
    };
    for (uint32_t k = 0; k < E; ++k) {
        if (!edge_src[k].empty()) {
            std::vector<uint32_t> keep;
            const SourceDef& s0 = sources[edge_src[k].front()];
            uint32_t surf0 = s0.surface_label.empty() ? 0 : s0.surface_label.front();
            for (uint32_t j : edge_src[k]) {
                const SourceDef& sj = sources[j];
                uint32_t sf = sj.surface_label.empty() ? 0 : sj.surface_label.front();
                if (sj.st == s0.st && sf == surf0) keep.push_back(j);
            }
            edge_src[k] = keep;
        }
        const uint32_t p0 = en[2 * (size_t)k], p1 = en[2 * (size_t)k + 1];
        const uint32_t rec0 = cls_node[p0], rec1 = cls_node[p1];
        bool treated = false;
        if (edge_bcs[k] != 0 && bc_type(edge_bcs[k]) == "pec") {
            b.edge_class[k] = 4;
            treated = true;
        } else if (!edge_src[k].empty()) {
            for (uint32_t s : edge_src[k])
                if (sources[s].st == "e") { b.edge_class[k] = 4; treated = true; break; }
        }
        if (!treated) {
            if (rec0 > 2 || rec1 > 2 || rec0 + rec1 == 1 || rec0 + rec1 == 3 || dvp_sig[p0] != dvp_sig[p1]) {
                b.edge_class[k] = 3;
                fracture_star(p0);
                fracture_star(p1);
            } else if (rec0 == 2 || rec1 == 2 || rec0 == 0 || rec1 == 0) {
                b.edge_class[k] = 1;
                b.h_row[k] = H_size++;
            } else {
                b.edge_class[k] = 2;
                p_index[k] = P_size++;
            }
        } else {
            bnd_index[k] = B_size++;
        }
        if (b.edge_class[k] == 4 && edge_bcs[k] != 0) {
            // ComputeEdgeBC: BCs[BndToBC[edge_bcs[e]]].GetValue() (D:3899-3905)
            uint32_t l = edge_bcs[k] < m.bnd_to_bc.size() ? m.bnd_to_bc[edge_bcs[k]] : 0;
            b.fixed_value[k] = bcs[l].val;
        }
    }
    // face indices of the fractional magnetic unknowns (D:6501-6541): n_index
    // (class 1), r_index (class 2) and the offset of each fractured tet's
    // block of F_c (primal_is_fractured), handed out in face order
    std::vector<int32_t> n_index(F, -1), r_index(F, -1), tet_off(T, -1);
    int32_t N_size = 0, R_size = 0, S_size = 0;
    for (uint32_t f = 0; f < F; ++f) {
        if (face_class[f] == 1) n_index[f] = N_size++;
        else if (face_class[f] == 2) r_index[f] = R_size++;
        else
            for (int q = 0; q < ftv_n[f]; ++q) {
                const int32_t t = ftv_t[(size_t)f * 2 + q];
                if (tet_off[t] < 0) { tet_off[t] = S_size; S_size += 4; }
            }
    }
    // lossy or magnetic-loss media: dga steps the fractional unknowns
    const bool frac = is_dga && (P_size > 0 || Q_size > 0 || R_size > 0 || S_size > 0);

    // ---- C (faces x edges) and its transpose
    {
        CsrStore& C = b.C;
        C.nrows = (int32_t)F;
        C.ncols = (int32_t)E;
        C.rowptr.resize(F + 1);
        C.col.resize((size_t)F * 3);
        C.sgn.resize((size_t)F * 3);
        for (uint32_t f = 0; f < F; ++f) {
            C.rowptr[f] = (int64_t)f * 3;
            for (int q = 0; q < 3; ++q) {
                C.col[(size_t)f * 3 + q] = fte[(size_t)f * 3 + q];
                C.sgn[(size_t)f * 3 + q] = q == 1 ? -1 : 1;
            }
        }
        C.rowptr[F] = (int64_t)F * 3;
        CsrStore& Ct = b.Ct;
        Ct.nrows = (int32_t)E;
        Ct.ncols = (int32_t)F;
        Ct.rowptr.assign(E + 1, 0);
        for (size_t q = 0; q < C.col.size(); ++q) Ct.rowptr[C.col[q] + 1]++;
        for (uint32_t e = 0; e < E; ++e) Ct.rowptr[e + 1] += Ct.rowptr[e];
        Ct.col.resize(C.col.size());
        Ct.sgn.resize(C.col.size());
        std::vector<int64_t> pos(Ct.rowptr.begin(), Ct.rowptr.end() - 1);
        for (uint32_t f = 0; f < F; ++f)  // increasing face order within each edge row
            for (int q = 0; q < 3; ++q) {
                int32_t e = C.col[(size_t)f * 3 + q];
                Ct.col[pos[e]] = (int32_t)f;
                Ct.sgn[pos[e]++] = C.sgn[(size_t)f * 3 + q];
            }
    }
    b.bnd_edge = bnd_edges;

    clk.mark("classes, C");
    // ---- local geometry per tet
    // local edge order after sort_unique + swap(1,2) + swap(1,3) (D:10314-10316):
    // 0:(n0n1) 1:(n1n2) 2:(n0n2) 3:(n0n3) 4:(n1n3) 5:(n2n3)
    const int kLocEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    auto edge_id = [&](uint32_t a, uint32_t c) -> int32_t {
        // a < c; nte of a is sorted by edge id and edges are sorted by (a,c)
        for (uint32_t q = nte_ptr[a]; q < nte_ptr[a + 1]; ++q) {
            uint32_t e = nte[q];
            if (en[2 * (size_t)e] == a && en[2 * (size_t)e + 1] == c) return (int32_t)e;
        }
        throw Error("internal: edge lookup");
    };
    std::vector<int32_t> tet_edges((size_t)T * 6);
    for (uint32_t t = 0; t < T; ++t) {
        const int32_t* n = &b.vols[(size_t)t * 4];
        for (int q = 0; q < 6; ++q) tet_edges[(size_t)t * 6 + q] = edge_id(n[kLocEdge[q][0]], n[kLocEdge[q][1]]);
    }
    // dual_area_vectors (D:10301-10351)
    auto dual_area_vectors = [&](uint32_t t, Vec3 out[6]) {
        Vec3 ebs[6], fbs[4];
        for (int q = 0; q < 6; ++q) ebs[q] = edge_bars[tet_edges[(size_t)t * 6 + q]];
        for (int q = 0; q < 4; ++q) fbs[q] = face_bars[b.vtf[(size_t)t * 4 + q]];
        Vec3 vb = vol_bar(t);
        double sign = b.cell_volumes[t] < 0 ? -0.5 : 0.5;
        out[0] = cross(vb - ebs[0], fbs[2] - fbs[3]) * sign;
        out[1] = cross(vb - ebs[1], fbs[0] - fbs[3]) * sign;
        out[2] = cross(vb - ebs[2], fbs[3] - fbs[1]) * sign;
        out[3] = cross(vb - ebs[3], fbs[1] - fbs[2]) * sign;
        out[4] = cross(vb - ebs[4], fbs[2] - fbs[0]) * sign;
        out[5] = cross(vb - ebs[5], fbs[0] - fbs[1]) * sign;
    };
    // primal_area_vectors (D:10397-10418)
    auto primal_area_vectors = [&](uint32_t t, Vec3 out[4]) {
        const int32_t* n = &b.vols[(size_t)t * 4];
        Vec3 ev0 = P(n[1]) - P(n[0]), ev1 = P(n[2]) - P(n[1]), ev2 = P(n[2]) - P(n[0]);
        Vec3 ev3 = P(n[3]) - P(n[0]), ev4 = P(n[3]) - P(n[1]);
        out[0] = cross(ev1, ev4) / 2.0;
        out[1] = cross(ev2, ev3) / 2.0;
        out[2] = cross(ev0, ev3) / 2.0;
        out[3] = cross(ev0, ev2) / 2.0;
    };
    // local_M of the Codecasa / gfdtd mu block (D:8757-8776)
    // with w = chi instead of mu: local_Z of the magnetic losses (D:8778-8801)
    auto local_W = [&](uint32_t t, const Mat3& mu, double M[16]) {
        Vec3 fv[6];
        dual_area_vectors(t, fv);
        auto d = [&](int i, int j) { return dot(fv[i], mu * fv[j]); };
        double coeff = 36.0 / std::fabs(b.cell_volumes[t]);
        M[0] = (d(0, 0) + d(2, 2) + d(3, 3)) * coeff;
        M[5] = (d(0, 0) + d(1, 1) + d(4, 4)) * coeff;
        M[10] = (d(1, 1) + d(2, 2) + d(5, 5)) * coeff;
        M[15] = (d(3, 3) + d(4, 4) + d(5, 5)) * coeff;
        M[1] = (-d(1, 2) - d(3, 4)) * coeff;
        M[2] = (-d(0, 1) + d(3, 5)) * coeff;
        M[3] = (d(0, 4) + d(2, 5)) * coeff;
        M[6] = (-d(0, 2) - d(4, 5)) * coeff;
        M[7] = (d(0, 3) - d(1, 5)) * coeff;
        M[11] = (-d(2, 3) - d(1, 4)) * coeff;
        M[4] = M[1]; M[8] = M[2]; M[12] = M[3]; M[9] = M[6]; M[13] = M[7]; M[14] = M[11];
    };
    // (map::at: the lambdas below also run on host threads)
    auto local_M = [&](uint32_t t, double M[16]) { local_W(t, materials.at(b.vol_material[t]).mu, M); };

    clk.mark("tet edges");
    // ---- N = sum over tets of local_M^-1 (rows n_index = face id for class-1
    // faces). The per-tet inverses on all threads, then each face row sums
    // its (at most two) tets' rows in increasing tet order: the order of the
    // sequential loop over tets.
    {
        std::vector<double> ninv((size_t)T * 16);
        std::atomic<bool> bad{false};
#pragma omp parallel
        {
            std::vector<double> Mv(16), inv;
            double M[16];
#pragma omp for schedule(static)
            for (int64_t t = 0; t < (int64_t)T; ++t) {
                local_M((uint32_t)t, M);
                Mv.assign(M, M + 16);
                if (!spd_inverse(Mv, 4, inv)) { bad = true; continue; }
                std::copy(inv.begin(), inv.end(), ninv.begin() + t * 16);
            }
        }
        if (bad) throw Error("local mu block not positive definite (degenerate tet)");
        b.N = assemble_rows((int32_t)F, (int32_t)F, [&](int32_t f, std::vector<RowEntry>& out) {
            for (int q = 0; q < ftv_n[f]; ++q) {
                const int32_t t = ftv_t[(size_t)f * 2 + q];
                const int32_t* tf = &b.vtf[(size_t)t * 4];
                int j = 0;
                while (tf[j] != f) ++j;
                for (int k = 0; k < 4; ++k) out.push_back({tf[k], ninv[(size_t)t * 16 + j * 4 + k]});
            }
        });
    }

    clk.mark("N");
    // Corner data of tet t at its local node k (D:8913-9078): local edges and
    // the face area vector / sign factor that define the dual basis f_i.
    // k=0: edges (0,2,3) faces (1,2,3) s=-1 ; k=1: (0,1,4) faces (0,2,3) s=(+,-,-)
    // k=2: (1,2,5) faces (1,0,3) s=(+,+,-) ; k=3: (3,4,5) faces (0,1,2) s=+
    const int kCornerEdges[4][3] = {{0, 2, 3}, {0, 1, 4}, {1, 2, 5}, {3, 4, 5}};
    const int kCornerFaces[4][3] = {{1, 2, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};
    const double kCornerSign[4][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {1, 1, 1}};

    const double sim_dt_request = sim.step;
    double t_step = 0;
    set_output(m, sim, b);

    if (is_dga) {
        // ---- node-star eps blocks, local inverses and H (D:8819-9121, 9263-9371)
        // Rows of H: the class-1 edges. With lossy media (fractional
        // unknowns) this H serves the dt estimate only, and every
        // non-Dirichlet edge is a row: the lossless operator of the same
        // mesh (ComputeDGATimeStep works on the lossless inverses,
        // D:9088-9141); the stepped operators are assembled once dt is known.
        std::vector<int32_t> est_row = b.h_row;
        int32_t est_size = H_size;
        if (frac) {
            est_size = 0;
            for (uint32_t e = 0; e < E; ++e) est_row[e] = b.edge_class[e] != 4 ? est_size++ : -1;
        }
        // Node-star corner term of tet t at its local node k: the dual basis
        // f_i of the corner's three edges (D:8913-9078).
        auto star_corner = [&](uint32_t t, int k, Vec3 f[3]) {
            Vec3 fa[4];
            primal_area_vectors(t, fa);
            const double vol = std::fabs(b.cell_volumes[t]);
            for (int i = 0; i < 3; ++i) {
                int face = kCornerFaces[k][i];
                double sg = vtf_sgn[(size_t)t * 4 + face];
                f[i] = fa[face] * (kCornerSign[k][i] * 2.0 / 3.0 / vol * sg);
            }
            return vol;
        };
        // E: the node-star eps mass blocks summed without the class-4
        // identity rows (E_fracs assembled into E_trip, D:9435-9445), the
        // matrix of the energy diagnostic We = 1/2 U.E U (D:2773). Row e =
        // (a, b) collects the corners at a, then at b, of the tets holding e
        // (each in associated-volume order): per entry the order of the
        // reference's loop over node stars.
        b.E = assemble_rows((int32_t)E, (int32_t)E, [&](int32_t row, std::vector<RowEntry>& out) {
            for (int s = 0; s < 2; ++s) {
                const uint32_t nid = en[2 * (size_t)row + s], other = en[2 * (size_t)row + 1 - s];
                for (uint32_t q = av_cnt[nid]; q < av_cnt[nid + 1]; ++q) {
                    const uint32_t t = av[q];
                    const int32_t* n = &b.vols[(size_t)t * 4];
                    if (n[0] != (int32_t)other && n[1] != (int32_t)other && n[2] != (int32_t)other &&
                        n[3] != (int32_t)other)
                        continue;
                    int k = 0;
                    while (n[k] != (int32_t)nid) ++k;
                    Vec3 f[3];
                    const double vol = star_corner(t, k, f);
                    const Mat3& eps = materials.at(b.vol_material[t]).epsilon;
                    int32_t e[3];
                    int i0 = 0;
                    for (int i = 0; i < 3; ++i) {
                        e[i] = tet_edges[(size_t)t * 6 + kCornerEdges[k][i]];
                        if (e[i] == row) i0 = i;
                    }
                    for (int j = 0; j < 3; ++j) out.push_back({e[j], dot(f[i0], eps * f[j]) * vol / 4});
                }
            }
        });
        // Node-star blocks LE, their inverses LH (class-4 rows/cols set to the
        // identity for the inversion and zeroed after, D:9270-9333), one star
        // per task; each kept at lh_off[node] (deg x deg)
        std::vector<int64_t> lh_off(np + 1, 0);
        for (uint32_t nid = 0; nid < np; ++nid) {
            const int64_t deg = nte_ptr[nid + 1] - nte_ptr[nid];
            lh_off[nid + 1] = lh_off[nid] + deg * deg;
        }
        std::vector<double> LHall((size_t)lh_off[np]);
        std::atomic<bool> bad{false};
#pragma omp parallel
        {
            std::vector<double> LE, LH;
#pragma omp for schedule(dynamic, 256)
            for (int64_t nid = 0; nid < (int64_t)np; ++nid) {
                const uint32_t deg = nte_ptr[nid + 1] - nte_ptr[nid];
                if (deg == 0) continue;
                const uint32_t* gi = &nte[nte_ptr[nid]];
                LE.assign((size_t)deg * deg, 0.0);
                auto loc = [&](int32_t e) { return (uint32_t)(std::lower_bound(gi, gi + deg, (uint32_t)e) - gi); };
                for (uint32_t q = av_cnt[nid]; q < av_cnt[nid + 1]; ++q) {
                    const uint32_t t = av[q];
                    const int32_t* n = &b.vols[(size_t)t * 4];
                    int k = 0;
                    while (n[k] != (int32_t)nid) ++k;
                    Vec3 f[3];
                    const double vol = star_corner(t, k, f);
                    const Mat3& eps = materials.at(b.vol_material[t]).epsilon;
                    uint32_t le[3];
                    for (int i = 0; i < 3; ++i) le[i] = loc(tet_edges[(size_t)t * 6 + kCornerEdges[k][i]]);
                    for (int i = 0; i < 3; ++i) LE[(size_t)le[i] * deg + le[i]] += dot(f[i], eps * f[i]) * vol / 4;
                    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
                    for (auto& pq : pairs) {
                        double v = dot(f[pq[0]], eps * f[pq[1]]) * vol / 4;
                        double& a = LE[(size_t)le[pq[0]] * deg + le[pq[1]]];
                        a = a + v;
                        LE[(size_t)le[pq[1]] * deg + le[pq[0]]] = a;
                    }
                }
                for (uint32_t j = 0; j < deg; ++j)
                    if (b.edge_class[gi[j]] == 4) {
                        for (uint32_t c = 0; c < deg; ++c) LE[(size_t)j * deg + c] = LE[(size_t)c * deg + j] = 0;
                        LE[(size_t)j * deg + j] = 1;
                    }
                if (!spd_inverse(LE, (int)deg, LH)) { bad = true; continue; }
                for (uint32_t j = 0; j < deg; ++j)
                    if (b.edge_class[gi[j]] == 4)
                        for (uint32_t c = 0; c < deg; ++c) LH[(size_t)j * deg + c] = LH[(size_t)c * deg + j] = 0;
                std::copy(LH.begin(), LH.end(), LHall.begin() + lh_off[nid]);
            }
        }
        if (bad) throw Error("node-star eps block not positive definite");
        // H (rows: est_row; D:9345-9371 pushes each star's upper triangle and
        // its mirror, so entry (j, k) of a star reads LH[min(j,k), max(j,k)]).
        // Row e = (a, b) sums the stars of a, then b, over the non-class-4
        // edges of each: the row pattern star(a) U star(b) without class 4.
        std::vector<int32_t> est_edge(est_size);
        for (uint32_t e = 0; e < E; ++e)
            if (est_row[e] >= 0) est_edge[est_row[e]] = (int32_t)e;
        b.H = assemble_rows(est_size, (int32_t)E, [&](int32_t r, std::vector<RowEntry>& out) {
            const int32_t row = est_edge[r];
            for (int s = 0; s < 2; ++s) {
                const uint32_t nid = en[2 * (size_t)row + s];
                const uint32_t deg = nte_ptr[nid + 1] - nte_ptr[nid];
                const uint32_t* gi = &nte[nte_ptr[nid]];
                const uint32_t jj = (uint32_t)(std::lower_bound(gi, gi + deg, (uint32_t)row) - gi);
                const double* LH = &LHall[lh_off[nid]];
                for (uint32_t kk = 0; kk < deg; ++kk) {
                    if (b.edge_class[gi[kk]] == 4) continue;
                    const uint32_t lo = std::min(jj, kk), hi = std::max(jj, kk);
                    out.push_back({(int32_t)gi[kk], LH[(size_t)lo * deg + hi]});
                }
            }
        });
        std::vector<double>().swap(LHall);
    clk.mark("H (node stars)");
        // ---- dt: power iteration on H * Ct * N * C (ComputeDGATimeStep D:7121)
        {
            std::vector<double> U(E), cu(F), Fw(F), cf(E);
            auto apply = [&](const std::vector<double>& x, std::vector<double>& y) {
                // x, y live on H rows
#pragma omp parallel for schedule(static)
                for (int64_t e = 0; e < (int64_t)E; ++e) U[e] = est_row[e] >= 0 ? x[est_row[e]] : 0.0;
#pragma omp parallel for schedule(static)
                for (int64_t f = 0; f < (int64_t)F; ++f) {
                    double s = 0;
                    for (int q = 0; q < 3; ++q) s += b.C.sgn[(size_t)f * 3 + q] * U[b.C.col[(size_t)f * 3 + q]];
                    cu[f] = s;
                }
#pragma omp parallel for schedule(static)
                for (int64_t f = 0; f < (int64_t)F; ++f) {
                    double s = 0;
                    for (int64_t q = b.N.rowptr[f]; q < b.N.rowptr[f + 1]; ++q) s += b.N.val[q] * cu[b.N.col[q]];
                    Fw[f] = s;
                }
#pragma omp parallel for schedule(static)
                for (int64_t e = 0; e < (int64_t)E; ++e) {
                    double s = 0;
                    for (int64_t q = b.Ct.rowptr[e]; q < b.Ct.rowptr[e + 1]; ++q) s += b.Ct.sgn[q] * Fw[b.Ct.col[q]];
                    cf[e] = s;
                }
#pragma omp parallel for schedule(static)
                for (int32_t r = 0; r < est_size; ++r) {
                    double s = 0;
                    for (int64_t q = b.H.rowptr[r]; q < b.H.rowptr[r + 1]; ++q) s += b.H.val[q] * cf[b.H.col[q]];
                    y[r] = s;
                }
            };
            double lam = power_lambda_max((size_t)est_size, apply, 60) * 1.02;
            b.stable_dt = lam > 0 ? 2.0 / std::sqrt(lam) : 0;
        }
    } else if (is_fem) {
        // ---- fem: Whitney element mass E and SigMat (ConstructFEMaterialMatrices
        // D:7917-8058). Per tet, the Whitney functions of its six edges at
        // the four nodes (weight 1/40) and at the barycentres of the four
        // faces (weight 9/40) of the degree-3 rule, |V| sum w_k . eps w_j,
        // pushed for j <= k (and mirrored) in the reference's order.
        Triplets Etr, Str;
        bool lossy = false;
        for (uint32_t t = 0; t < T; ++t) {
            const Material& mt = materials[b.vol_material[t]];
            if (mt.chi.det() != 0)
                throw Error("method fem with magnetic losses (chi) is not supported: the reference scales its "
                            "local Z by a time step it has not computed yet (D:7798, D:8074)");
            const bool lossy_t = mt.sigma.det() != 0;
            lossy = lossy || lossy_t;
            const int32_t* n = &b.vols[(size_t)t * 4];
            // barycentric coordinates: lambda_i(p) = [p 1] . wn[:, i]
            double M4[16], wn[16];
            for (int i = 0; i < 4; ++i) {
                const Vec3& q = P(n[i]);
                M4[i * 4 + 0] = q.x; M4[i * 4 + 1] = q.y; M4[i * 4 + 2] = q.z; M4[i * 4 + 3] = 1.0;
            }
            {
                std::vector<double> A(M4, M4 + 16), inv;
                if (!general_inverse(A, 4, inv)) throw Error("degenerate tetrahedron in the fem assembly");
                for (int z = 0; z < 16; ++z) wn[z] = inv[z];
            }
            Vec3 g[4];
            for (int i = 0; i < 4; ++i) g[i] = Vec3(wn[0 * 4 + i], wn[1 * 4 + i], wn[2 * 4 + i]);
            auto whitney = [&](const Vec3& q, Vec3 w[6]) {
                double l[4];
                for (int i = 0; i < 4; ++i) l[i] = q.x * wn[i] + q.y * wn[4 + i] + q.z * wn[8 + i] + wn[12 + i];
                const int le[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
                for (int k = 0; k < 6; ++k) w[k] = l[le[k][0]] * g[le[k][1]] - l[le[k][1]] * g[le[k][0]];
            };
            double sn[6][6] = {}, sf[6][6] = {}, sns[6][6] = {}, sfs[6][6] = {};
            for (int h = 0; h < 4; ++h) {
                Vec3 w1[6], w2[6];
                whitney(P(n[h]), w1);
                whitney(face_bars[b.vtf[(size_t)t * 4 + h]], w2);
                for (int j = 0; j < 6; ++j)
                    for (int k = j; k < 6; ++k) {
                        sn[j][k] = sn[j][k] + dot(w1[k], mt.epsilon * w1[j]);
                        sf[j][k] = sf[j][k] + dot(w2[k], mt.epsilon * w2[j]);
                        sns[j][k] = sns[j][k] + dot(w1[k], mt.sigma * w1[j]);
                        sfs[j][k] = sfs[j][k] + dot(w2[k], mt.sigma * w2[j]);
                    }
            }
            const int32_t* te = &tet_edges[(size_t)t * 6];
            const double vol = std::fabs(b.cell_volumes[t]);
            for (int j = 0; j < 6; ++j)
                for (int k = j; k < 6; ++k) {
                    const double val = vol * (sn[j][k] + sf[j][k] * 9.0) / 40.0;
                    if (val != 0) {
                        Etr.add(te[j], te[k], val);
                        if (k != j) Etr.add(te[k], te[j], val);
                    }
                    if (lossy_t) {
                        const double sv = vol * (sns[j][k] + sfs[j][k] * 9.0) / 40.0;
                        if (sv != 0) {
                            Str.add(te[j], te[k], sv);
                            if (k != j) Str.add(te[k], te[j], sv);
                        }
                    }
                }
        }
        b.E = Etr.to_csr((int32_t)E, (int32_t)E);
        if (lossy) {
            // SigMat on E's pattern (its entries are a subset)
            CsrStore S = Str.to_csr((int32_t)E, (int32_t)E);
            b.Sig = b.E;
            std::fill(b.Sig.val.begin(), b.Sig.val.end(), 0.0);
            for (uint32_t r = 0; r < E; ++r)
                for (int64_t q = S.rowptr[r]; q < S.rowptr[r + 1]; ++q) csr_at(b.Sig, (int32_t)r, S.col[q]) = S.val[q];
        }
    clk.mark("E, SigMat (Whitney elements)");
    }
    if (!is_dga) {
        if (!is_fem) {
        // ---- gfdtd: E from the tet-corner dual basis (D:8379-8420), A, RHSmat1.
        // Row e = (a, b) collects the corners at a, then at b, of the tets
        // holding e (each in associated-volume order): per entry the order of
        // the reference's loop over nodes and their tets. The corner's 3x3
        // block is pushed in the order of D:8420-8429, which touches each
        // entry once.
        // SigMat: the same corner terms with sigma, from the tets whose sigma
        // tensor is invertible (D:8600-8612); kept on E's pattern (the other
        // corners enter it as zeros)
        bool lossy = false;
        for (uint32_t t = 0; t < T && !lossy; ++t) lossy = materials[b.vol_material[t]].sigma.det() != 0;
        auto corner_rows = [&](bool sigma) {
            return [&, sigma](int32_t row, std::vector<RowEntry>& out) {
                for (int s = 0; s < 2; ++s) {
                    const uint32_t nid = en[2 * (size_t)row + s], other = en[2 * (size_t)row + 1 - s];
                    for (uint32_t q = av_cnt[nid]; q < av_cnt[nid + 1]; ++q) {
                        const uint32_t t = av[q];
                        const int32_t* n = &b.vols[(size_t)t * 4];
                        if (n[0] != (int32_t)other && n[1] != (int32_t)other && n[2] != (int32_t)other &&
                            n[3] != (int32_t)other)
                            continue;
                        int k = 0;
                        while (n[k] != (int32_t)nid) ++k;
                        const Material& mt = materials.at(b.vol_material[t]);
                        const bool use = !sigma || mt.sigma.det() != 0;
                        int32_t e[3];
                        int i0 = 0;
                        for (int i = 0; i < 3; ++i) {
                            e[i] = tet_edges[(size_t)t * 6 + kCornerEdges[k][i]];
                            if (e[i] == row) i0 = i;
                        }
                        if (!use) {
                            for (int j = 0; j < 3; ++j) out.push_back({e[j], 0.0});
                            continue;
                        }
                        const double omega = std::fabs(b.cell_volumes[t]) / double(4);
                        const Mat3& w = sigma ? mt.sigma : mt.epsilon;
                        Vec3 a1 = edge_bars[e[0]] - P(nid), a2 = edge_bars[e[1]] - P(nid), a3 = edge_bars[e[2]] - P(nid);
                        double nc = std::pow(dot(a1, cross(a2, a3)), -1);
                        Vec3 f[3] = {nc * cross(a2, a3), nc * cross(a3, a1), nc * cross(a1, a2)};
                        Mat3 ow;
                        for (int z = 0; z < 9; ++z) ow.m[z] = omega * w.m[z];
                        for (int j = 0; j < 3; ++j) out.push_back({e[j], dot(f[i0], ow * f[j])});
                    }
                }
            };
        };
        b.E = assemble_rows((int32_t)E, (int32_t)E, corner_rows(false));
        if (lossy) b.Sig = assemble_rows((int32_t)E, (int32_t)E, corner_rows(true));
    clk.mark("E (corner terms)");
        }
        // free (non-Dirichlet) edges: compressed_dirichlet (D:8145-8171;
        // fem: D:7703-7729, the same rule)
        std::vector<int32_t> is_dir(E, -1);
        for (uint32_t i = 0; i < E; ++i) {
            bool no_bc = edge_bcs[i] == 0 || bc_type(edge_bcs[i]) == "none";
            if (!no_bc) continue;
            bool dirich = false;
            for (uint32_t s : edge_src[i])
                if (sources[s].st == "e") { dirich = true; break; }
            if (!dirich) {
                is_dir[i] = (int32_t)b.free_edges.size();
                b.free_edges.push_back((int32_t)i);
            }
        }
        const int32_t nfree = (int32_t)b.free_edges.size();
        // A = E restricted to free edges; RHSmat1 = E(free rows, dirichlet cols);
        // with lossy media both take E + dt/2 SigMat once dt is known (below)
        b.A.nrows = b.A.ncols = nfree;
        b.A.rowptr.assign(nfree + 1, 0);
        b.RHS1.nrows = b.RHS1.ncols = (int32_t)E;
        b.RHS1.rowptr.assign(E + 1, 0);
        for (uint32_t r = 0; r < E; ++r) {
            if (is_dir[r] >= 0) {
                for (int64_t q = b.E.rowptr[r]; q < b.E.rowptr[r + 1]; ++q) {
                    int32_t c = b.E.col[q];
                    if (is_dir[c] >= 0) { b.A.col.push_back(is_dir[c]); b.A.val.push_back(b.E.val[q]); }
                    else { b.RHS1.col.push_back(c); b.RHS1.val.push_back(b.E.val[q]); }
                }
                b.A.rowptr[is_dir[r] + 1] = (int64_t)b.A.col.size();
            }
            b.RHS1.rowptr[r + 1] = (int64_t)b.RHS1.col.size();
        }
    clk.mark("A, RHSmat1");
        // ---- dt: generalized power iteration on (Ct N C, A) (ComputeFEMTimeStep D:6928)
        {
            std::vector<double> U(E), cu(F), Fw(F), cf(E), r(nfree), z(nfree), pp(nfree), qq(nfree), y(nfree);
            std::vector<double> dinv(nfree, 1.0);
            for (int32_t i = 0; i < nfree; ++i)
                for (int64_t q = b.A.rowptr[i]; q < b.A.rowptr[i + 1]; ++q)
                    if (b.A.col[q] == i && b.A.val[q] != 0) dinv[i] = 1.0 / b.A.val[q];
            auto Kx = [&](const std::vector<double>& x, std::vector<double>& out) {
#pragma omp parallel for schedule(static)
                for (int64_t e = 0; e < (int64_t)E; ++e) U[e] = 0.0;
#pragma omp parallel for schedule(static)
                for (int32_t i = 0; i < nfree; ++i) U[b.free_edges[i]] = x[i];
#pragma omp parallel for schedule(static)
                for (int64_t f = 0; f < (int64_t)F; ++f) {
                    double s = 0;
                    for (int q = 0; q < 3; ++q) s += b.C.sgn[(size_t)f * 3 + q] * U[b.C.col[(size_t)f * 3 + q]];
                    cu[f] = s;
                }
#pragma omp parallel for schedule(static)
                for (int64_t f = 0; f < (int64_t)F; ++f) {
                    double s = 0;
                    for (int64_t q = b.N.rowptr[f]; q < b.N.rowptr[f + 1]; ++q) s += b.N.val[q] * cu[b.N.col[q]];
                    Fw[f] = s;
                }
#pragma omp parallel for schedule(static)
                for (int32_t i = 0; i < nfree; ++i) {
                    int32_t e = b.free_edges[i];
                    double s = 0;
                    for (int64_t q = b.Ct.rowptr[e]; q < b.Ct.rowptr[e + 1]; ++q) s += b.Ct.sgn[q] * Fw[b.Ct.col[q]];
                    out[i] = s;
                }
            };
            auto Ax = [&](const std::vector<double>& x, std::vector<double>& out) {
#pragma omp parallel for schedule(static)
                for (int32_t i = 0; i < nfree; ++i) {
                    double s = 0;
                    for (int64_t q = b.A.rowptr[i]; q < b.A.rowptr[i + 1]; ++q) s += b.A.val[q] * x[b.A.col[q]];
                    out[i] = s;
                }
            };
            auto dotv = [&](const std::vector<double>& a, const std::vector<double>& c) {
                return par_dot(a.data(), c.data(), a.size());
            };
            // x enters as the start (the previous power step's output: once
            // the iteration settles, y_k = A^-1 K x_k is close to y_{k-1})
            // and the tolerance is relative to ||rhs||, the zero start's
            // residual, so a warm start ends in fewer iterations at the same
            // accuracy
            auto solveA = [&](const std::vector<double>& rhs, std::vector<double>& x) {
                Ax(x, qq);
#pragma omp parallel for schedule(static)
                for (int32_t i = 0; i < nfree; ++i) r[i] = rhs[i] - qq[i];
                double r0 = std::sqrt(dotv(rhs, rhs)), rho = 0, rho_p;
                if (r0 == 0) {
                    std::fill(x.begin(), x.end(), 0.0);
                    return;
                }
                if (std::sqrt(dotv(r, r)) / r0 < 1e-6) return;
                for (int it = 0; it < 200; ++it) {
#pragma omp parallel for schedule(static)
                    for (int32_t i = 0; i < nfree; ++i) z[i] = dinv[i] * r[i];
                    rho_p = rho;
                    rho = dotv(r, z);
                    const double beta = it == 0 ? 0.0 : rho / rho_p;
#pragma omp parallel for schedule(static)
                    for (int32_t i = 0; i < nfree; ++i) pp[i] = it == 0 ? z[i] : z[i] + beta * pp[i];
                    Ax(pp, qq);
                    double a = rho / dotv(pp, qq);
#pragma omp parallel for schedule(static)
                    for (int32_t i = 0; i < nfree; ++i) { x[i] += a * pp[i]; r[i] -= a * qq[i]; }
                    if (std::sqrt(dotv(r, r)) / r0 < 1e-6) break;
                }
            };
            std::vector<double> kx(nfree), w(nfree);
            // Rayleigh quotient in the A inner product after power steps
            std::vector<double> x(nfree);
            uint64_t st = 0x9E3779B97F4A7C15ull;
            for (auto& v : x) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; v = (double)(st >> 11) * 1.1102230246251565e-16 - 0.5; }
            double lam = 0;
            // K x of the iterate is kept from the Rayleigh quotient of the
            // step before, so each step applies K once
            Kx(x, kx);
            for (int it = 0; it < 25; ++it) {
                solveA(kx, y);
                double n2 = std::sqrt(dotv(y, y));
                if (n2 == 0) break;
#pragma omp parallel for schedule(static)
                for (int32_t i = 0; i < nfree; ++i) x[i] = y[i] / n2;
                Kx(x, kx);
                Ax(x, w);
                lam = dotv(x, kx) / dotv(x, w);
            }
            lam *= 1.02;
            b.stable_dt = lam > 0 ? 2.0 / std::sqrt(lam) : 0;
        }
        b.cg_tol = sim.tol;
    }
    t_step = sim.courant * b.stable_dt;
    if (sim_dt_request != 0) {
        if (sim_dt_request > t_step) b.possibly_unstable = true;
        else t_step = sim_dt_request;
    }
    b.prob.dt = t_step;
    if (frac) {
        // ---- lossy / magnetic-loss dga: the fractional operators once dt is
        // known (ConstructCodecasaMaterialMatrices, D:9180-9481). Triplets are
        // pushed in the reference's order and summed (overwritten for M and
        // T) in that order.
        const double hdt = 0.5 * t_step;
        // tets: local_N = (M + dt/2 Z)^-1 and local_RS = local_N (M - dt/2 Z)
        // where chi is invertible, else M^-1 and the identity (D:9180-9257)
        Triplets Ntr, Trtr, Str, Tstr, Ttr;
        std::vector<double> Rr(R_size, 0.0);
        {
            std::vector<double> A(16), LN;
            double Mb[16], Zb[16], RS[16];
            for (uint32_t t = 0; t < T; ++t) {
                const Mat3& chi = materials[b.vol_material[t]].chi;
                local_M(t, Mb);
                const bool lossy_t = chi.det() != 0;
                if (lossy_t) {
                    local_W(t, chi, Zb);
                    for (int z = 0; z < 16; ++z) A[z] = Mb[z] + hdt * Zb[z];
                } else {
                    A.assign(Mb, Mb + 16);
                }
                if (!spd_inverse(A, 4, LN)) throw Error("local mu block not positive definite (degenerate tet)");
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j) {
                        double v = 0;
                        if (lossy_t)
                            for (int q = 0; q < 4; ++q) v += LN[i * 4 + q] * (Mb[q * 4 + j] - hdt * Zb[q * 4 + j]);
                        else
                            v = i == j ? 1.0 : 0.0;
                        RS[i * 4 + j] = v;
                    }
                const int32_t* fid = &b.vtf[(size_t)t * 4];  // abs(vtf_list[t]): f4, f3, f2, f1
                const int32_t off = tet_off[t];
                for (int jj = 0; jj < 4; ++jj) {
                    const int32_t j = fid[jj];
                    if (face_class[j] == 1) {
                        for (int kk = 0; kk < 4; ++kk)
                            if (LN[jj * 4 + kk] != 0) Ntr.add(n_index[j], fid[kk], LN[jj * 4 + kk]);
                        Ttr.add(j, n_index[j], 1.0);
                    } else if (face_class[j] == 2) {
                        Rr[r_index[j]] = RS[jj * 4 + jj];
                        for (int kk = 0; kk < 4; ++kk)
                            if (LN[jj * 4 + kk] != 0) Trtr.add(r_index[j], fid[kk], LN[jj * 4 + kk]);
                        Ttr.add(j, N_size + r_index[j], 1.0);
                    } else {
                        if (off < 0) throw Error("internal: class-3 face of an unfractured tet");
                        Ttr.add(j, N_size + R_size + off + jj, 1.0);
                    }
                    if (off >= 0)
                        for (int kk = jj; kk < 4; ++kk) {
                            const double rs = RS[jj * 4 + kk], ln = LN[jj * 4 + kk];
                            if (rs != 0) {
                                Str.add(off + jj, off + kk, rs);
                                if (jj != kk) Str.add(off + kk, off + jj, rs);
                            }
                            if (ln != 0) {
                                Tstr.add(off + jj, fid[kk], ln);
                                if (jj != kk) Tstr.add(off + kk, fid[jj], ln);
                            }
                        }
                }
            }
        }
        // node stars: local_E and local_S (D:8913-9078), class-4 rows and
        // columns set to the identity (E) and zero (S), local_H = (E +
        // dt/2 S)^-1 and local_PQ = local_H (E - dt/2 S) at stars touching a
        // tet with invertible sigma, else E^-1 and the identity; class-4 rows
        // and columns of local_H zeroed, of local_PQ zeroed with a unit
        // diagonal (D:9262-9333). The mirrored H and Mp entries go to the row
        // of the partner edge's h_index / p_index, which is 0 when the partner
        // belongs to another class (reproduced; D:9363-9366, D:9382-9385).
        Triplets Htr, Mptr, Mqtr, Qtr, Mtr;
        std::vector<double> Pp(P_size, 0.0);
        {
            std::vector<double> LE, LS, A, LH, LPQ;
            Vec3 fa[4];
            const int32_t Uf_h = H_size, Uf_q = H_size + P_size, Uf_b = H_size + P_size + Q_size;
            auto hidx = [&](uint32_t e) { return b.edge_class[e] == 1 ? b.h_row[e] : 0; };
            auto pidx = [&](uint32_t e) { return b.edge_class[e] == 2 ? p_index[e] : 0; };
            for (uint32_t nid = 0; nid < np; ++nid) {
                const uint32_t deg = nte_ptr[nid + 1] - nte_ptr[nid];
                if (deg == 0) continue;
                const uint32_t* gi = &nte[nte_ptr[nid]];
                LE.assign((size_t)deg * deg, 0.0);
                LS.assign((size_t)deg * deg, 0.0);
                bool sigma_node = false;
                auto loc = [&](int32_t e) { return (uint32_t)(std::lower_bound(gi, gi + deg, (uint32_t)e) - gi); };
                for (uint32_t q = av_cnt[nid]; q < av_cnt[nid + 1]; ++q) {
                    const uint32_t t = av[q];
                    const int32_t* n = &b.vols[(size_t)t * 4];
                    int k = 0;
                    while (n[k] != (int32_t)nid) ++k;
                    primal_area_vectors(t, fa);
                    const double vol = std::fabs(b.cell_volumes[t]);
                    const Mat3& eps = materials[b.vol_material[t]].epsilon;
                    const Mat3& sig = materials[b.vol_material[t]].sigma;
                    const bool lossy_t = sig.det() != 0;
                    if (lossy_t) sigma_node = true;
                    Vec3 f[3];
                    uint32_t le[3];
                    for (int i = 0; i < 3; ++i) {
                        int face = kCornerFaces[k][i];
                        double sg = vtf_sgn[(size_t)t * 4 + face];
                        f[i] = fa[face] * (kCornerSign[k][i] * 2.0 / 3.0 / vol * sg);
                        le[i] = loc(tet_edges[(size_t)t * 6 + kCornerEdges[k][i]]);
                    }
                    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
                    for (std::vector<double>* M : {&LE, &LS}) {
                        const bool s_part = M == &LS;
                        if (s_part && !lossy_t) continue;
                        const Mat3& w = s_part ? sig : eps;
                        std::vector<double>& L = *M;
                        for (int i = 0; i < 3; ++i) L[(size_t)le[i] * deg + le[i]] += dot(f[i], w * f[i]) * vol / 4;
                        for (auto& pq : pairs) {
                            double v = dot(f[pq[0]], w * f[pq[1]]) * vol / 4;
                            double& a = L[(size_t)le[pq[0]] * deg + le[pq[1]]];
                            a = a + v;
                            L[(size_t)le[pq[1]] * deg + le[pq[0]]] = a;
                        }
                    }
                }
                for (uint32_t j = 0; j < deg; ++j)
                    if (b.edge_class[gi[j]] == 4) {
                        for (uint32_t c = 0; c < deg; ++c) {
                            LE[(size_t)j * deg + c] = LE[(size_t)c * deg + j] = 0;
                            if (sigma_node) LS[(size_t)j * deg + c] = LS[(size_t)c * deg + j] = 0;
                        }
                        LE[(size_t)j * deg + j] = 1;
                    }
                A = LE;
                if (sigma_node)
                    for (size_t z = 0; z < A.size(); ++z) A[z] += hdt * LS[z];
                if (!spd_inverse(A, (int)deg, LH)) throw Error("node-star eps block not positive definite");
                LPQ.assign((size_t)deg * deg, 0.0);
                if (sigma_node) {
                    for (uint32_t i = 0; i < deg; ++i)
                        for (uint32_t j = 0; j < deg; ++j) {
                            double v = 0;
                            for (uint32_t q = 0; q < deg; ++q)
                                v += LH[(size_t)i * deg + q] * (LE[(size_t)q * deg + j] - hdt * LS[(size_t)q * deg + j]);
                            LPQ[(size_t)i * deg + j] = v;
                        }
                } else {
                    for (uint32_t i = 0; i < deg; ++i) LPQ[(size_t)i * deg + i] = 1.0;
                }
                for (uint32_t j = 0; j < deg; ++j)
                    if (b.edge_class[gi[j]] == 4)
                        for (uint32_t c = 0; c < deg; ++c) {
                            LH[(size_t)j * deg + c] = LH[(size_t)c * deg + j] = 0;
                            LPQ[(size_t)j * deg + c] = LPQ[(size_t)c * deg + j] = 0;
                            LPQ[(size_t)j * deg + j] = 1;
                        }
                const int32_t off = star_off[nid];
                for (uint32_t jj = 0; jj < deg; ++jj) {
                    const uint32_t j = gi[jj];
                    const int cls = b.edge_class[j];
                    if (cls == 1 || cls == 2) {
                        Triplets& tr = cls == 1 ? Htr : Mptr;
                        const int32_t row = cls == 1 ? b.h_row[j] : p_index[j];
                        const int32_t nrows = cls == 1 ? H_size : P_size;
                        if (cls == 2) Pp[p_index[j]] = LPQ[(size_t)jj * deg + jj];
                        for (uint32_t kk = jj; kk < deg; ++kk) {
                            const double v = LH[(size_t)jj * deg + kk];
                            if (b.edge_class[gi[kk]] == 4 || v == 0) continue;
                            tr.add(row, (int32_t)gi[kk], v);
                            if (jj != kk && nrows > 0) tr.add(cls == 1 ? hidx(gi[kk]) : pidx(gi[kk]), (int32_t)j, v);
                        }
                        Mtr.add((int32_t)j, cls == 1 ? row : Uf_h + row, 1.0);
                    } else if (cls == 3) {
                        if (off < 0) throw Error("internal: class-3 edge of an unfractured node star");
                        Mtr.add((int32_t)j, Uf_q + off + (int32_t)jj, 1.0);
                    } else {
                        Mtr.add((int32_t)j, Uf_b + bnd_index[j], 1.0);
                    }
                    if (off >= 0 && cls != 4)
                        for (uint32_t kk = 0; kk < deg; ++kk) {
                            if (b.edge_class[gi[kk]] == 4) continue;
                            const double pq = LPQ[(size_t)jj * deg + kk], h = LH[(size_t)jj * deg + kk];
                            if (pq != 0) Qtr.add(off + (int32_t)jj, off + (int32_t)kk, pq);
                            if (h != 0) Mqtr.add(off + (int32_t)jj, (int32_t)gi[kk], h);
                        }
                }
            }
        }
        // Stacked for the stepping (rows in fractional order):
        //   Ku = [H; Mp; Mq; 0 (boundary rows)], Ru = [I; diag P_p; Q; I]
        //   Kf = [N; Tr; Ts],                   Rf = [I; diag R_r; S]
        // so that U_frac <- Ru U_frac_old + dt Ku curl_f and F_frac <- Rf
        // F_frac_old - dt Kf curl_u, U = M U_frac, F = T F_frac.
        const int32_t nU = H_size + P_size + Q_size + B_size, nF = N_size + R_size + S_size;
        auto stack = [](std::initializer_list<const CsrStore*> parts, int32_t ncols, int32_t nrows_total) {
            CsrStore m;
            m.nrows = nrows_total;
            m.ncols = ncols;
            m.rowptr.assign(1, 0);
            for (const CsrStore* p : parts) {
                for (int32_t r = 0; r < p->nrows; ++r) {
                    for (int64_t q = p->rowptr[r]; q < p->rowptr[r + 1]; ++q) {
                        m.col.push_back(p->col[q]);
                        m.val.push_back(p->val[q]);
                    }
                    m.rowptr.push_back((int64_t)m.col.size());
                }
            }
            while ((int32_t)m.rowptr.size() < nrows_total + 1) m.rowptr.push_back((int64_t)m.col.size());
            return m;
        };
        CsrStore Hm = Htr.to_csr(H_size, (int32_t)E), Mp = Mptr.to_csr(P_size, (int32_t)E),
                 Mq = Mqtr.to_csr(Q_size, (int32_t)E);
        b.frac_Ku = stack({&Hm, &Mp, &Mq}, (int32_t)E, nU);
        CsrStore Qm = Qtr.to_csr(Q_size, Q_size);
        {
            CsrStore& R = b.frac_Ru;
            R.nrows = R.ncols = nU;
            R.rowptr.assign(1, 0);
            for (int32_t r = 0; r < nU; ++r) {
                if (r < H_size || r >= H_size + P_size + Q_size) {
                    R.col.push_back(r);
                    R.val.push_back(1.0);
                } else if (r < H_size + P_size) {
                    R.col.push_back(r);
                    R.val.push_back(Pp[r - H_size]);
                } else {
                    const int32_t qr = r - H_size - P_size;
                    for (int64_t q = Qm.rowptr[qr]; q < Qm.rowptr[qr + 1]; ++q) {
                        R.col.push_back(H_size + P_size + Qm.col[q]);
                        R.val.push_back(Qm.val[q]);
                    }
                }
                R.rowptr.push_back((int64_t)R.col.size());
            }
        }
        CsrStore Nm = Ntr.to_csr(N_size, (int32_t)F), Tr = Trtr.to_csr(R_size, (int32_t)F),
                 Ts = Tstr.to_csr(S_size, (int32_t)F);
        b.frac_Kf = stack({&Nm, &Tr, &Ts}, (int32_t)F, nF);
        CsrStore Sm = Str.to_csr(S_size, S_size);
        {
            CsrStore& R = b.frac_Rf;
            R.nrows = R.ncols = nF;
            R.rowptr.assign(1, 0);
            for (int32_t r = 0; r < nF; ++r) {
                if (r < N_size) {
                    R.col.push_back(r);
                    R.val.push_back(1.0);
                } else if (r < N_size + R_size) {
                    R.col.push_back(r);
                    R.val.push_back(Rr[r - N_size]);
                } else {
                    const int32_t sr = r - N_size - R_size;
                    for (int64_t q = Sm.rowptr[sr]; q < Sm.rowptr[sr + 1]; ++q) {
                        R.col.push_back(N_size + R_size + Sm.col[q]);
                        R.val.push_back(Sm.val[q]);
                    }
                }
                R.rowptr.push_back((int64_t)R.col.size());
            }
        }
        b.frac_M = Mtr.to_csr((int32_t)E, nU, true);
        b.frac_T = Ttr.to_csr((int32_t)F, nF, true);
        // the U_frac entry a boundary value or source writes each step: U_d
        // for class-4 edges, U_a for class-1 edges (a d source resets it,
        // D:843-849); U_b and U_c are overwritten from their old values by
        // the update, so writes to them have no effect (D:888-890)
        b.frac_slot.assign(E, -1);
        for (uint32_t e = 0; e < E; ++e) {
            if (b.edge_class[e] == 1) b.frac_slot[e] = b.h_row[e];
            else if (b.edge_class[e] == 4) b.frac_slot[e] = H_size + P_size + Q_size + bnd_index[e];
        }
        b.frac_sizes[0] = H_size; b.frac_sizes[1] = P_size; b.frac_sizes[2] = Q_size; b.frac_sizes[3] = B_size;
        b.frac_sizes[4] = N_size; b.frac_sizes[5] = R_size; b.frac_sizes[6] = S_size;
        // the lossless H above was the dt estimate's; the stepping uses Ku
        b.H = CsrStore();
    }
    if (!is_dga && (is_fem || !b.Sig.rowptr.empty())) {
        // gfdtd: SysMat = E + (dt/2) SigMat (D:8646-8668); fem: SysMat =
        // E (1/dt^2) + (0.5/dt) SigMat (D:8086-8090); split into A (free x
        // free) and RHSmat1 (free rows, Dirichlet columns); SigMat has E's
        // pattern
        const double h = 0.5 * t_step, inv2 = 1 / (t_step * t_step), sh = 0.5 / t_step;
        const bool sig = !b.Sig.rowptr.empty();
        std::vector<uint8_t> is_free(E, 0);
        for (int32_t e : b.free_edges) is_free[e] = 1;
        size_t qa = 0, q1 = 0;
        for (uint32_t r = 0; r < E; ++r) {
            if (!is_free[r]) continue;
            for (int64_t q = b.E.rowptr[r]; q < b.E.rowptr[r + 1]; ++q) {
                const double v = is_fem ? (sig ? b.E.val[q] * inv2 + sh * b.Sig.val[q] : b.E.val[q] * inv2)
                                        : b.E.val[q] + h * b.Sig.val[q];
                if (is_free[b.E.col[q]]) b.A.val[qa++] = v;
                else b.RHS1.val[q1++] = v;
            }
        }
    }

    clk.mark("dt estimate, coefficients");
    // ---- probes (FindProbe D:3645) with Whitney E and face H weights
    uint32_t root = 0;
    auto det4 = [](const Vec3 r[4]) {
        // det of rows [x y z 1]
        Vec3 a = r[1] - r[0], c = r[2] - r[0], d = r[3] - r[0];
        return -dot(a, cross(c, d));
    };
    auto find_probe = [&](const Vec3& pv) -> uint32_t {
        const double tol = 1e-12;
        std::vector<uint8_t> colour(T, 0);
        std::vector<uint32_t> queue;
        queue.push_back(root);
        colour[0]++;
        for (size_t k = 0; k < queue.size(); ++k) {
            uint32_t v = queue[k];
            const int32_t* n = &b.vols[(size_t)v * 4];
            Vec3 r[4] = {P(n[0]), P(n[1]), P(n[2]), P(n[3])};
            double d0 = det4(r);
            bool ok = true;
            for (int s = 0; s < 4 && ok; ++s) {
                Vec3 rr[4] = {r[0], r[1], r[2], r[3]};
                rr[s] = pv;
                double d1 = det4(rr);
                ok = (d0 >= -tol && d1 >= -tol) || (d0 <= tol && d1 <= tol);
            }
            if (ok) { root = v; return v; }
            for (int q = 0; q < 4; ++q) {
                int32_t f = b.vtf[(size_t)v * 4 + q];
                uint32_t cn = (uint32_t)ftv_t[(size_t)f * 2];
                if (cn == v) cn = (uint32_t)ftv_t[(size_t)f * 2 + ftv_n[f] - 1];
                if (colour[cn] == 0) { queue.push_back(cn); colour[cn]++; }
            }
            colour[v]++;
        }
        return T;
    };
    if (b.out_mode == 1 || b.out_mode == 2) {
        const OutputDef& o = m.outputs.at(sim.output);
        for (const Vec3& pv : o.probes) {
            uint32_t t = find_probe(pv);
            if (t >= T) continue;
            const int32_t* n = &b.vols[(size_t)t * 4];
            Vec3 r[4] = {P(n[0]), P(n[1]), P(n[2]), P(n[3])};
            // barycentric coordinates and gradients: lambda_i(p) with grad
            // lambda_i = rows of the inverse of [r1-r0, r2-r0, r3-r0]
            Vec3 a = r[1] - r[0], c = r[2] - r[0], d = r[3] - r[0];
            double det = dot(a, cross(c, d));
            Vec3 g1 = cross(c, d) / det, g2 = cross(d, a) / det, g3 = cross(a, c) / det;
            Vec3 g0 = -(g1 + g2 + g3);
            Vec3 grad[4] = {g0, g1, g2, g3};
            double lam[4];
            Vec3 dp = pv - r[0];
            lam[1] = dot(g1, dp); lam[2] = dot(g2, dp); lam[3] = dot(g3, dp);
            lam[0] = 1 - lam[1] - lam[2] - lam[3];
            tf_probe pr{};
            tf_probe_geom g{};
            pr.ne = 6;
            pr.nh = 4;
            for (int q = 0; q < 6; ++q) {
                int i = kLocEdge[q][0], j = kLocEdge[q][1];
                Vec3 w = lam[i] * grad[j] - lam[j] * grad[i];
                pr.e_idx[q] = tet_edges[(size_t)t * 6 + q];
                pr.e_w[q][0] = w.x; pr.e_w[q][1] = w.y; pr.e_w[q][2] = w.z;
                g.edges[q] = pr.e_idx[q];
            }
            // GetMagneticField D:3877-3897
            Vec3 pev[6];
            for (int q = 0; q < 6; ++q) pev[q] = P(n[kLocEdge[q][1]]) - P(n[kLocEdge[q][0]]);
            Vec3 coef[4] = {pev[0] + pev[2] + pev[3], pev[0] - pev[1] - pev[4], -pev[2] - pev[1] + pev[5],
                            pev[3] + pev[4] + pev[5]};
            Mat3 mui = materials[b.vol_material[t]].mu.inverse();
            double v3 = 3 * b.cell_volumes[t];
            for (int q = 0; q < 4; ++q) {
                Vec3 w = (mui * (coef[q] / v3)) * 0.25;
                pr.h_idx[q] = b.vtf[(size_t)t * 4 + q];
                pr.h_w[q][0] = w.x; pr.h_w[q][1] = w.y; pr.h_w[q][2] = w.z;
                g.faces[q] = pr.h_idx[q];
            }
            g.elem = (int32_t)t;
            g.point[0] = pv.x; g.point[1] = pv.y; g.point[2] = pv.z;
            for (int q = 0; q < 4; ++q) { g.nodes[q][0] = r[q].x; g.nodes[q][1] = r[q].y; g.nodes[q][2] = r[q].z; }
            std::memcpy(g.mu, materials[b.vol_material[t]].mu.m, sizeof(g.mu));
            g.volume = b.cell_volumes[t];
            b.probes.push_back(pr);
            b.probe_geom.push_back(g);
            b.probe_points.push_back(pv.x); b.probe_points.push_back(pv.y); b.probe_points.push_back(pv.z);
        }
        // j dipoles (D:429-462, placed only with probe output as in the
        // reference): the edge nearest to the location, by barycentre, among
        // the edges around the first node of the tet that holds it, in
        // increasing edge id (the first of equal distances wins)
        for (uint32_t sl : sim.sources) {
            const SourceDef& s = sources[sl];
            if (s.st != "j") continue;
            if (is_dga)
                throw Error("j dipole sources in dga are not supported by this build (D:835-865 reads an uninitialised value)");
            const Vec3 loc(s.location[0], s.location[1], s.location[2]);
            const uint32_t t = find_probe(loc);
            if (t >= T) continue;  // out of the mesh: the reference skips it with a warning
            const uint32_t node = (uint32_t)b.vols[(size_t)t * 4];
            int64_t best = -1;
            double bd = 0;
            for (uint32_t q = nte_ptr[node]; q < nte_ptr[node + 1]; ++q) {
                const double dq = norm(edge_bars[nte[q]] - loc);
                if (best < 0 || dq < bd) { bd = dq; best = nte[q]; }
            }
            if (best >= 0) edge_src[best].push_back(sl);
        }
    }

    clk.mark("probes");
    // ---- sources on edges: h (bnd dual vector), e (edge vector)
    std::map<uint32_t, int32_t> src_index;
    for (uint32_t sl : sim.sources) {
        src_index[sl] = (int32_t)b.sources.size();
        b.sources.push_back(make_source(sources[sl]));
        b.src_desc.push_back(describe_source(sources[sl]));
    }
    for (uint32_t e = 0; e < E; ++e) {
        if (edge_src[e].empty()) continue;
        const SourceDef& s0 = sources[edge_src[e].front()];
        tf_source_edge d{};
        d.edge = (int32_t)e;
        d.field = make_source(s0).field;
        if (d.field == TF_FIELD_H && bnd_edges[e] < 0) continue;  // gfdtd checks bnd_edges >= 0 (D:1488)
        // an edge whose first source is neither h nor j goes to
        // ComputeEfieldSource (D:827-842, D:1497-1499), which sums the "e"
        // sources only (D:3928): for a d edge that is U = 0. gfdtd overwrites
        // such a free edge with the solve (no effect); dga resets its U_a to
        // 0 before the H update, carried here by zero-weight terms
        if (d.field == TF_FIELD_D && !is_dga) continue;
        if (edge_src[e].size() > 4)
            throw Error("more than 4 sources on one edge (edge " + std::to_string(e) + ")");
        Vec3 ev = P(en[2 * (size_t)e + 1]) - P(en[2 * (size_t)e]);
        Vec3 vec = d.field == TF_FIELD_H ? bnd_dual[bnd_edges[e]] : ev;
        const Vec3& eb = edge_bars[e];
        d.bary[0] = eb.x; d.bary[1] = eb.y; d.bary[2] = eb.z;
        d.vec[0] = vec.x; d.vec[1] = vec.y; d.vec[2] = vec.z;
        for (uint32_t sl : edge_src[e]) {
            const SourceDef& s = sources[sl];
            d.src[d.n_src++] = src_index[sl];
            if (make_source(s).field != d.field) continue;
            double w = 0.0;
            if (d.field != TF_FIELD_D) {
                Vec3 sp = SourceEval::space(s, eb);
                w = dot(sp, vec);
                if (d.field == TF_FIELD_J) w = (1.0 / norm(ev)) * w;
            }
            b.terms.push_back(tf_source_term{(int32_t)e, src_index[sl], w});
        }
        b.src_edge_desc.push_back(d);
    }

    {
        const Lattice lat = mesh_lattice(pts, en, E);
        b.edge_rank = locality_rank(lat, pts, en, 2, E);
        b.face_rank = locality_rank(lat, pts, fn, 3, F);
        if (b.edge_rank.empty() || b.face_rank.empty()) { b.edge_rank.clear(); b.face_rank.clear(); }
    }
    b.face_nodes.assign(fn.begin(), fn.end());
    b.edge_nodes.assign(en.begin(), en.end());
    b.pts.resize(np * 3);
    for (uint64_t i = 0; i < np; ++i) { b.pts[i * 3] = pts[i].x; b.pts[i * 3 + 1] = pts[i].y; b.pts[i * 3 + 2] = pts[i].z; }
    b.n_pts = (int32_t)np;
    b.prob.method = is_dga ? TF_METHOD_DGA : is_fem ? TF_METHOD_FEM : TF_METHOD_GFDTD;
    b.prob.n_faces = (int32_t)F;
    b.prob.n_vols = (int32_t)T;
    b.prob.duration = sim.duration;
    b.first_step = 1;
    // for (i=1; i*t_step <= simulation_time; ++i)  (D:810, D:1459)
    int64_t i = 1;
    while ((double)i * t_step <= sim.duration) ++i;
    b.n_steps = i - 1;
    clk.mark("sources");
}

}  // namespace tfh
