// tetfit - command-line driver of the MI355X build, the counterpart of the
// reference's dga_main.cpp:38-67 (parse the .fdtd input, then run every
// simulation it defines): each simulation is meshed and assembled by
// libtetfit_host, stepped on the GPU by libtetfit_hip, and its outputs are
// written in the reference's text formats:
//   <name><sim:05d>_probe_E.dat, <name><sim:05d>_probe_H.dat  (probepoint
//       mode, D:1948-2030: x y z t Ex Ey Ez per probe and sample, setw(15);
//       H at t - dt/2; cylindrical frame converted as at D:2032-2074)
//   poynting.dat  (t, Poynting flux, losses per step, D:945-950, D:1573-1578,
//       D:1936-1941; losses from tf_read_losses: gfdtd's SigMat terms, 0 for
//       dga (lossless here) and fdtd (Joule weights only in l2norm mode))
//   <name><sim:05d>_energy_electric.dat, _energy_magnetic.dat  (l2norm and
//       maxerror modes, D:2250-2282: t, W, W_analytic, W_difference, and the
//       trapezoid time integrals of the three, D:2795-2818; H at t + dt/2).
//       The analytic comparison is not built: the analytic field is taken as
//       zero, so W_analytic and its integral are 0 and the difference
//       columns repeat the numeric ones.
//
// usage: tetfit input.fdtd [--fp32] [--learned] [--tol X]
//   --fp32     fp32 fields and operators (fp64 reductions)
//   --learned  gfdtd: learned preconditioner (trained on the first 200 solves)
//   --tol X    CG tolerance instead of the simulation's (gfdtd)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "tetfit_hip.h"
#include "tetfit_host.h"

namespace {

[[noreturn]] void die(const std::string& what) {
    std::cerr << "tetfit: " << what << std::endl;
    std::exit(1);
}

void check(tf_status rc, const char* call) {
    if (rc != TF_OK) die(std::string(call) + ": " + tf_last_error());
}

std::string file_name(const std::string& name, int32_t sim, const char* suffix) {
    std::ostringstream ss;
    ss << name << std::setw(5) << std::setfill('0') << sim << suffix;
    return ss.str();
}

// One probe line: three coordinates, the time and three components, each in a
// field of 15 with the stream's default precision (D:1972-1988).
void write_line(std::ofstream& os, const double c[3], double t, const double v[3]) {
    os << std::setw(15) << c[0] << " " << std::setw(15) << c[1] << " " << std::setw(15) << c[2] << " "
       << std::setw(15) << t << " " << std::setw(15) << v[0] << " " << std::setw(15) << v[1] << " "
       << std::setw(15) << v[2] << std::endl;
}

// The two energy files of D:2250-2282 from the sampled (We, Wh) series;
// the time integrals are the reference's trapezoid sums over the sample
// times (D:2795-2808), 0 at the first sample.
void write_energy(const std::string& name, int32_t sim, const std::vector<double>& t,
                  const std::vector<double>& vals, double dt) {
    std::ostringstream se, sh;
    se << name << std::setw(5) << std::setfill('0') << sim << "_energy_electric.dat";
    sh << name << std::setw(5) << std::setfill('0') << sim << "_energy_magnetic.dat";
    std::ofstream os(se.str()), osh(sh.str());
    double cum_e = 0, cum_h = 0;
    for (size_t k = 0; k < t.size(); ++k) {
        const double we = vals[2 * k], wh = vals[2 * k + 1];
        if (k > 0) {
            cum_e += 0.5 * (t[k] - t[k - 1]) * (vals[2 * k - 2] + we);
            cum_h += 0.5 * (t[k] - t[k - 1]) * (vals[2 * k - 1] + wh);
        }
        // columns: t, W, W_analytic (0), W_difference (= W), and their integrals
        os << std::setw(15) << t[k] << " " << std::setw(15) << we << " " << std::setw(15) << 0.0 << " "
           << std::setw(15) << we << " " << std::setw(15) << cum_e << " " << std::setw(15) << 0.0 << " "
           << std::setw(15) << cum_e << std::endl;
        osh << std::setw(15) << t[k] + 0.5 * dt << " " << std::setw(15) << wh << " " << std::setw(15) << 0.0 << " "
            << std::setw(15) << wh << " " << std::setw(15) << cum_h << " " << std::setw(15) << 0.0 << " "
            << std::setw(15) << cum_h << std::endl;
    }
}

void write_probes(const std::string& name, int32_t sim, int32_t frame, const double* points, int32_t np,
                  const std::vector<double>& times, const std::vector<double>& vals, double dt) {
    std::ofstream os(file_name(name, sim, "_probe_E.dat")), osh(file_name(name, sim, "_probe_H.dat"));
    if (!os || !osh) die("cannot open the probe output files of " + name);
    for (size_t k = 0; k < times.size(); ++k)
        for (int32_t p = 0; p < np; ++p) {
            const double* pt = points + 3 * p;
            const double* v = &vals[(k * np + p) * 6];
            if (frame == 0) {
                write_line(os, pt, times[k], v);
                write_line(osh, pt, times[k] - 0.5 * dt, v + 3);
            } else if (frame == 1) {
                // cylindrical: (rho, theta, z) and the field rotated into (r, theta, z)
                const double rho = std::sqrt(pt[0] * pt[0] + pt[1] * pt[1]);
                double theta;
                if (pt[0] == 0 && pt[1] == 0) theta = 0;
                else if (pt[0] >= 0) theta = std::asin(pt[1] / rho);
                else theta = -std::asin(pt[1] / rho) + M_PI;
                double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 1}};
                if (rho > 0) {
                    m[0][0] = pt[0] / rho; m[0][1] = pt[1] / rho;
                    m[1][0] = -pt[1] / rho; m[1][1] = pt[0] / rho;
                }
                double e[3], h[3];
                for (int i = 0; i < 3; ++i) {
                    e[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
                    h[i] = m[i][0] * v[3] + m[i][1] * v[4] + m[i][2] * v[5];
                }
                const double c[3] = {rho, theta, pt[2]};
                write_line(os, c, times[k], e);
                write_line(osh, c, times[k] - 0.5 * dt, h);
            }
            // spherical frame: the reference writes no probe lines (D:1966-2075)
        }
}

}  // namespace

int main(int argc, char** argv) {
    std::cout << "    TetFIT time-domain Maxwell solver - MI355X stepping backend\n" << std::endl;
    std::string input;
    bool fp32 = false, learned = false;
    double tol_override = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--fp32") fp32 = true;
        else if (a == "--learned") learned = true;
        else if (a == "--tol" && i + 1 < argc) tol_override = std::atof(argv[++i]);
        else if (!a.empty() && a[0] != '-' && input.empty()) input = a;
        else die("usage: tetfit input.fdtd [--fp32] [--learned] [--tol X]");
    }
    if (input.empty()) die("usage: tetfit input.fdtd [--fp32] [--learned] [--tol X]");
    if (tf_device_count() < 1) die("no HIP device");

    tf_model* model = nullptr;
    if (tf_model_load(input.c_str(), &model) != 0) die(tfh_last_error());
    const int32_t n_sims = tf_model_n_simulations(model);
    for (int32_t s = 0; s < n_sims; ++s) {
        const int32_t label = tf_model_simulation_label(model, s);
        const auto t0 = std::chrono::steady_clock::now();
        tf_build* b = nullptr;
        if (tf_build_simulation(model, label, &b) != 0) die(tfh_last_error());
        const tf_problem* prob = tf_build_problem(b);
        const char* method = prob->method == TF_METHOD_FDTD ? "fdtd" : prob->method == TF_METHOD_DGA ? "dga" : "gfdtd";
        const double t_pre = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "Simulation " << label << ": method " << method << ", " << prob->n_edges << " edges, "
                  << prob->n_faces << " faces (" << t_pre << " s set-up)" << std::endl;
        if (tf_build_possibly_unstable(b))
            std::cout << "WARNING: the time step size chosen by the user is possibly unstable! "
                         "(Upper stable limit will be used instead...)" << std::endl;
        std::cout << std::setw(20) << "Simulation time: " << std::setw(20) << prob->duration << " sec" << std::endl;
        std::cout << std::setw(20) << "Time step: " << std::setw(20) << prob->dt << " sec" << std::endl;

        tf_ctx* ctx = nullptr;
        check(tf_create(prob, fp32 ? TF_PREC_FP32 : TF_PREC_FP64, &ctx), "tf_create");
        const tf_source* src;
        const tf_source_term* terms;
        int32_t n_src;
        int64_t n_terms;
        tf_build_sources(b, &src, &n_src, &terms, &n_terms);
        check(tf_set_sources(ctx, src, n_src, terms, n_terms), "tf_set_sources");
        const tf_probe* probes;
        const double* points;
        int32_t n_probes;
        tf_build_probes(b, &probes, &points, &n_probes);
        check(tf_set_probes(ctx, probes, n_probes), "tf_set_probes");
        const int32_t mode = tf_build_output_mode(b);
        if (mode == 1 || mode == 2 || mode == 3)
            check(tf_set_output(ctx, tf_build_output_delay(b), tf_build_output_period(b)), "tf_set_output");
        if (mode == 2 || mode == 3) check(tf_set_energy_output(ctx, 1), "tf_set_energy_output");
        tf_cg_opts cg{};
        cg.tol = tol_override > 0 ? tol_override : tf_build_cg_tol(b);
        cg.maxiter = 100;  // D:1388
        cg.precond = learned ? TF_PRECOND_LEARNED : TF_PRECOND_JACOBI;

        // the loop, in tenths for progress lines
        const int64_t n_steps = tf_build_n_steps(b), first = tf_build_first_step(b);
        double seconds = 0;
        int64_t done = 0, cg_iters = 0, not_converged = 0;
        while (done < n_steps) {
            const int64_t n = std::min<int64_t>(std::max<int64_t>(n_steps / 10, 1), n_steps - done);
            tf_stats part{};
            check(tf_run(ctx, first + done, n, prob->method == TF_METHOD_GFDTD ? &cg : nullptr, &part), "tf_run");
            done += n;
            seconds += part.seconds;
            cg_iters += part.cg_iters_total;
            not_converged += part.cg_not_converged;
            std::cout << "-----------Progress: " << std::setw(3) << 100 * done / std::max<int64_t>(n_steps, 1)
                      << "% done in " << std::setw(9) << seconds << "s, " << std::setw(8)
                      << seconds / std::max<int64_t>(done, 1) << " s/step-----------" << std::endl;
        }
        if (prob->method == TF_METHOD_GFDTD)
            std::cout << "CG iterations: " << cg_iters << " (" << (double)cg_iters / std::max<int64_t>(n_steps, 1)
                      << " per step), steps at maxiter: " << not_converged << std::endl;

        // probe files
        if (mode == 1) {
            int64_t ns = 0;
            check(tf_read_probes(ctx, nullptr, nullptr, 0, &ns), "tf_read_probes");
            std::vector<double> times(ns), vals((size_t)ns * n_probes * 6);
            check(tf_read_probes(ctx, times.data(), vals.data(), ns, &ns), "tf_read_probes");
            write_probes(tf_build_output_name(b), label, tf_build_output_frame(b), points, n_probes, times, vals,
                         prob->dt);
        }
        if (mode == 2 || mode == 3) {
            int64_t ns = 0;
            check(tf_read_energy(ctx, nullptr, nullptr, 0, &ns), "tf_read_energy");
            std::vector<double> times(ns), vals(2 * (size_t)ns);
            check(tf_read_energy(ctx, times.data(), vals.data(), ns, &ns), "tf_read_energy");
            write_energy(tf_build_output_name(b), label, times, vals, prob->dt);
        }
        // poynting.dat: step k at (k + 1/2) dt (fdtd, dga) or (k + 1) dt (gfdtd)
        {
            int64_t np = 0, nl = 0;
            check(tf_read_poynting(ctx, nullptr, 0, &np), "tf_read_poynting");
            std::vector<double> poy(np), loss(np, 0.0);
            check(tf_read_poynting(ctx, poy.data(), np, &np), "tf_read_poynting");
            check(tf_read_losses(ctx, loss.data(), np, &nl), "tf_read_losses");
            std::ofstream pf("poynting.dat");
            const double off = prob->method == TF_METHOD_GFDTD ? 1.0 : 0.5;
            for (int64_t k = 0; k < np; ++k) pf << double(k + off) * prob->dt << "\t" << poy[k] << "\t" << loss[k] << std::endl;
        }
        check(tf_destroy(ctx), "tf_destroy");
        tf_build_free(b);
    }
    tf_model_free(model);
    return 0;
}
