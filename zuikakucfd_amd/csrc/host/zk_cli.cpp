// zk_cli.cpp -- drop-in host executables for the reference's solver binaries.
//
// Built four times: zk_weno (-DZK_CLI_WENO) replaces solver_WENO, and
// zk_dg2/3/4 (-DZK_CLI_DG_P=1/2/3) replace solver_DG2/3/4 (Makefile:35-38,91).
// Command line, input files, dt modes, step loop, progress cadence and the
// result/<step>/U{0..3}_0.dat outputs follow src_WENO/solver.cpp:89-213 and
// src_DG/solver.cpp:152-332 for a single rank; the update itself runs in
// libzk.so on the GPU.  Options after the three positional arguments select
// what the reference fixes at compile time or does not offer:
//   --flux hllc|hll          numerical flux
//   --limiter persson|tvb|tvb-linear|none|learned   (DG) limiter, default persson (the
//                            reference's); tvb-linear: TVB without the P=3 face-average test
//   --M <value>              TVB constant, default 50
//   --precision fp64|fp32    (WENO) bulk arithmetic precision
//   --cfl <nu>               (WENO) auto/time modes use dt = nu*h/lambda instead of the
//                            reference's dt = lambda*h (src_WENO/solver.cpp:150,169)
//   --positivity             WENO: face positivity check with first-order fallback;
//                            DG: Zhang-Shu scaling also at the face points and corners
//   --slip                   (WENO) bctype-1 cells are mirror slip walls (zk_set_slip_walls);
//                            without it they are left alone, as the reference's WENO does
//   --recon weno5|muscl-minmod|muscl-vanleer|muscl-superbee
//                            (WENO) face reconstruction, default weno5 (the reference's
//                            compiled choice; MUSCL is its muscl_reconstruct, fluid.cpp:203)
//   --weights <file>         (DG P=2, learned limiter) network parameters, raw little-endian
//                            float32 (zk_mlp_param_count values); default
//                            <executable dir>/../data/troubled_cell_p2.f32
//   --band <lo> <hi>         (learned) score band deferred to the TVB test, default 0.2 0.8
//   --audit <fraction>       (learned) random share of cells also checked by TVB, default 0.02
//   --ranks <n>              write U<c>_<r>.dat for r = 0..n-1 as `mpirun -np n` would
//                            (column slices with 3 exchanged ghost columns per inner side)
//   --output-every <n>       output interval in steps, default 100; 0 = the last step only
//   --dump                   also write result/<step>/state.f64 (all modes, full precision)
//   --quiet                  no per-step progress lines
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <initializer_list>
#include <iostream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "zk.h"

namespace fs = std::filesystem;

namespace {

struct Input {
    int nx = 0, ny = 0;
    double da = 0, gamma = 0;
    std::vector<double> rho, u, v, p;
    std::vector<int32_t> bct;
};

// Whitespace-separated values read rows first, as `file >> mat(i, j)`
// (src_WENO/fluid.cpp:33-49); a missing file is an error.
template <typename T>
bool read_matrix(const std::string &path, std::vector<T> &out, size_t n, std::string &err)
{
    std::ifstream f(path);
    if (!f) {
        err = "Cannot open file: " + path;
        return false;
    }
    out.resize(n);
    for (size_t k = 0; k < n; ++k) {
        if (!(f >> out[k])) {
            err = path + ": expected " + std::to_string(n) + " values";
            return false;
        }
    }
    return true;
}

bool read_input(const std::string &dir, Input &in, std::string &err)
{
    std::ifstream pf(dir + "/params.txt");
    if (!pf) {
        err = "Cannot open params.txt";
        return false;
    }
    pf >> in.nx >> in.ny >> in.da >> in.gamma;
    if (!pf || in.nx < 7 || in.ny < 7) {
        err = "params.txt: expected `nx ny da gamma` with nx, ny >= 7";
        return false;
    }
    const size_t n = (size_t)in.nx * in.ny;
    return read_matrix(dir + "/rho.dat", in.rho, n, err) && read_matrix(dir + "/u.dat", in.u, n, err) &&
           read_matrix(dir + "/v.dat", in.v, n, err) && read_matrix(dir + "/p.dat", in.p, n, err) &&
           read_matrix(dir + "/bctype.dat", in.bct, n, err);
}

// Eigen 3 default operator<< text: stream precision 6 (%g), every coefficient
// right-aligned to the widest one, single spaces, newline between rows, no
// trailing newline.
// Column slices [lo, hi) of the per-rank output of a `mpirun -np n` run:
// rank widths w_k = remain / (n - k) (splitMeshVertically,
// src_WENO/fluid.cpp:98-152), each rank's file holding its own columns plus
// 3 exchanged ghost columns on every inner side.  Interior values do not
// depend on the rank count, so slicing the whole-domain result reproduces them.
std::vector<std::pair<int, int>> rank_columns(int nx, int n)
{
    std::vector<std::pair<int, int>> out;
    int remain = nx, start = 0;
    for (int k = 0; k < n; ++k) {
        const int w = remain / (n - k);
        remain -= w;
        const int lg = k > 0 ? 3 : 0, rg = k < n - 1 ? 3 : 0;
        out.emplace_back(std::max(start - lg, 0), std::min(start + w + rg, nx));
        start += w;
    }
    return out;
}

void write_eigen(const fs::path &path, const double *a, int ny, int nx, int ld = -1)
{
    if (ld < 0) ld = nx;
    std::vector<std::string> s((size_t)ny * nx);
    size_t w = 0;
    char buf[64];
    for (size_t k = 0; k < s.size(); ++k) {
        std::snprintf(buf, sizeof(buf), "%g", a[(k / nx) * (size_t)ld + k % nx]);
        s[k] = buf;
        w = std::max(w, s[k].size());
    }
    std::string out;
    out.reserve(s.size() * (w + 1));
    for (int i = 0; i < ny; ++i) {
        if (i) out += '\n';
        for (int j = 0; j < nx; ++j) {
            if (j) out += ' ';
            const std::string &x = s[(size_t)i * nx + j];
            out.append(w - x.size(), ' ');
            out += x;
        }
    }
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot create " + path.string());
    f << out;
}

struct Opts {
    int flux = ZK_FLUX_HLLC;
    int limiter = ZK_LIMITER_PERSSON_ZS;
    double M = 50.0;
    int precision = ZK_PREC_FP64;
    double cfl = 0.0;  // 0: reference dt rule
    bool dump = false;
    bool quiet = false;
    bool positivity = false;
    int recon = ZK_RECON_WENO5;
    bool slip = false;
    std::string weights;
    int ranks = 1;  // per-rank output files as a `mpirun -np ranks` run would write
    int every = 100;  // output interval in steps (the reference's 100); 0: the last step only
    float band_lo = 0.2f, band_hi = 0.8f, audit = 0.02f;
};

bool parse_opts(int argc, char **argv, Opts &o)
{
    // an option value that is not one of the listed names is an error (exit 1)
    auto pick = [](const std::string &what, const std::string &v,
                   std::initializer_list<std::pair<const char *, int>> names, int &out) {
        for (const auto &n : names)
            if (v == n.first) {
                out = n.second;
                return true;
            }
        std::cerr << "unknown " << what << " " << v << "\n";
        return false;
    };
    try {
        for (int a = 4; a < argc; ++a) {
            std::string k = argv[a];
            auto val = [&](std::string &v) {
                if (a + 1 >= argc) return false;
                v = argv[++a];
                return true;
            };
            std::string v;
            if (k == "--flux" && val(v)) {
                if (!pick("flux", v, {{"hllc", ZK_FLUX_HLLC}, {"hll", ZK_FLUX_HLL}}, o.flux)) return false;
            } else if (k == "--limiter" && val(v)) {
                if (!pick("limiter", v,
                          {{"persson", ZK_LIMITER_PERSSON_ZS}, {"tvb", ZK_LIMITER_TVB}, {"none", ZK_LIMITER_NONE},
                           {"learned", ZK_LIMITER_LEARNED}, {"tvb-linear", ZK_LIMITER_TVB_LINEAR}},
                          o.limiter))
                    return false;
            } else if (k == "--M" && val(v)) o.M = std::stod(v);
            else if (k == "--precision" && val(v)) {
                if (!pick("precision", v, {{"fp64", ZK_PREC_FP64}, {"fp32", ZK_PREC_FP32_BULK}}, o.precision)) return false;
            } else if (k == "--cfl" && val(v)) o.cfl = std::stod(v);
            else if (k == "--weights" && val(v)) o.weights = v;
            else if (k == "--band" && a + 2 < argc) {
                o.band_lo = std::stof(argv[++a]);
                o.band_hi = std::stof(argv[++a]);
            } else if (k == "--audit" && val(v)) o.audit = std::stof(v);
            else if (k == "--ranks" && val(v)) o.ranks = std::max(1, std::stoi(v));
            else if (k == "--output-every" && val(v)) o.every = std::max(0, std::stoi(v));
            else if (k == "--dump") o.dump = true;
            else if (k == "--quiet") o.quiet = true;
            else if (k == "--positivity") o.positivity = true;
            else if (k == "--slip") o.slip = true;
            else if (k == "--recon" && val(v)) {
                if (!pick("reconstruction", v,
                          {{"weno5", ZK_RECON_WENO5}, {"muscl-minmod", ZK_RECON_MUSCL_MINMOD},
                           {"muscl-vanleer", ZK_RECON_MUSCL_VANLEER}, {"muscl-superbee", ZK_RECON_MUSCL_SUPERBEE}},
                          o.recon))
                    return false;
            } else {
                std::cerr << "unknown option " << k << "\n";
                return false;
            }
        }
    } catch (const std::exception &) {
        std::cerr << "bad numeric option value\n";
        return false;
    }
    return true;
}

// Loads the network of the learned indicator and attaches it to h; the
// network must outlive h's stepping (it is released at exit).
bool attach_network(zk_handle *h, const Opts &o, zk_mlp **out)
{
    std::string path = o.weights;
    if (path.empty()) {
        std::error_code ec;
        const fs::path exe = fs::read_symlink("/proc/self/exe", ec);
        path = (exe.parent_path().parent_path() / "data" / "troubled_cell_p2.f32").string();
    }
    std::vector<float> w(zk_mlp_param_count());
    std::ifstream f(path, std::ios::binary);
    if (!f || !f.read(reinterpret_cast<char *>(w.data()), w.size() * sizeof(float))) {
        std::cerr << "[zk] cannot read " << w.size() << " float32 network parameters from " << path << "\n";
        return false;
    }
    if (zk_mlp_create(0, 0, 0, out) != ZK_OK || zk_mlp_set_params(*out, w.data()) != ZK_OK) {
        std::cerr << "[zk] network: " << zk_mlp_last_error(*out) << "\n";
        return false;
    }
    if (zk_attach_mlp(h, *out, o.band_lo, o.band_hi, o.audit) != ZK_OK) {
        std::cerr << "[zk] zk_attach_mlp: " << zk_last_error(h) << "\n";
        return false;
    }
    return true;
}

int fail(zk_handle *h, const char *what)
{
    std::cerr << "[zk] " << what << ": " << zk_last_error(h) << "\n";
    if (h) zk_destroy(h);
    return 2;
}

void save(zk_handle *h, const Input &in, int nm, int step, bool dump, int ranks)
{
    const std::string folder = "result/" + std::to_string(step);
    fs::create_directories(folder);
    const size_t n = (size_t)in.nx * in.ny;
    std::vector<double> U(4 * n);
    zk_get_conserved(h, U.data(), U.data() + n, U.data() + 2 * n, U.data() + 3 * n);
    const auto cols = rank_columns(in.nx, ranks);
    for (int r = 0; r < ranks; ++r)
        for (int c = 0; c < 4; ++c)
            write_eigen(fs::path(folder) / ("U" + std::to_string(c) + "_" + std::to_string(r) + ".dat"),
                        U.data() + c * n + cols[r].first, in.ny, cols[r].second - cols[r].first, in.nx);
    if (dump) {
        std::vector<double> Q(4 * (size_t)nm * n);
        zk_get_state(h, Q.data());
        std::ofstream f(fs::path(folder) / "state.f64", std::ios::binary);
        const int32_t hdr[4] = {in.ny, in.nx, nm, 0};
        f.write(reinterpret_cast<const char *>(hdr), sizeof(hdr));
        f.write(reinterpret_cast<const char *>(Q.data()), Q.size() * sizeof(double));
    }
}

// End-of-run performance report with the quantities of the reference's
// printPerfReport (src_WENO/solver.cpp:19-86, src_DG/solver.cpp:86-150):
// wall time, mean time per step, I/O time and its share, and the per-rank
// compute table with its imbalance factor.  One process drives the whole grid
// here, so the table has one row and the imbalance is 1.  The mean per step
// is wall/steps for both methods (the WENO report divides by the rank count).
void perf_report(double wall_ms, double compute_ms, double io_ms, int steps, long long cells, const std::string &info)
{
    const std::string rule(40, '=');
    std::printf("\n%s\n  performance report\n%s\n", rule.c_str(), rule.c_str());
    std::printf("  total run time : %12.3f ms  (%.3f s)\n", wall_ms, wall_ms / 1000.0);
    std::printf("  mean per step  : %12.3f ms/step\n", steps > 0 ? wall_ms / steps : 0.0);
    std::printf("  max I/O time   : %12.3f ms  (%.1f%% of total)\n", io_ms, wall_ms > 0 ? io_ms / wall_ms * 100.0 : 0.0);
    std::printf("\n  -- load balance --\n  %6s %14s %10s %s\n  %s\n", "rank", "compute (ms)", "cells", " share",
                std::string(38, '-').c_str());
    std::printf("  %6d %14.3f %10lld  [||||||||||] 100.0%%\n", 0, compute_ms, cells);
    std::printf("  %s\n  max %.3f ms  min %.3f ms  mean %.3f ms\n", std::string(38, '-').c_str(), compute_ms, compute_ms,
                compute_ms);
    std::printf("  imbalance      : 1.000\n  efficiency     : 100.0%%\n  run            : %s\n%s\n", info.c_str(),
                rule.c_str());
}

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

}  // namespace

int main(int argc, char **argv)
{
    const auto wall0 = Clock::now();
    if (argc < 4) {
        std::cerr << "usage: " << argv[0] << " <mesh_dir> <dt|auto|time> <steps|t_end> [options]\n"
                  << "  e.g. " << argv[0] << " ./mesh 0.0001 1000   # fixed dt and step count\n"
                  << "       " << argv[0] << " ./mesh auto 2000     # CFL dt and step count\n"
                  << "       " << argv[0] << " ./mesh time 0.1      # CFL dt until t_end\n";
        return 1;
    }
    Opts o;
    if (!parse_opts(argc, argv, o)) return 1;
    const std::string mesh_dir = argv[1], dt_arg = argv[2], third = argv[3];
    const bool auto_dt = dt_arg == "auto", time_mode = dt_arg == "time";
    int timesteps = 0;
    double t_end = 0.0, dt_user = 0.0;
    try {
        if (time_mode) t_end = std::stod(third);
        else {
            timesteps = std::stoi(third);
            if (!auto_dt) dt_user = std::stod(dt_arg);
        }
    } catch (const std::exception &) {
        std::cerr << "bad numeric argument\n";
        return 1;
    }

    Input in;
    std::string err;
    if (!read_input(mesh_dir, in, err)) {
        std::cerr << err << "\n";
        return 1;
    }

    zk_config cfg{};
#ifdef ZK_CLI_WENO
    cfg.method = ZK_METHOD_WENO;
    cfg.P = 0;
    const int nm = 1;
#else
    cfg.method = ZK_METHOD_DG;
    cfg.P = ZK_CLI_DG_P;
    const int nm = (ZK_CLI_DG_P + 1) * (ZK_CLI_DG_P + 1);
#endif
    cfg.precision = o.precision;
    cfg.flux = o.flux;
    cfg.limiter = o.limiter;
    cfg.tvb_M = o.M;
    cfg.device = 0;
    cfg.positivity = o.positivity ? 1 : 0;
    zk_handle *h = nullptr;
    if (zk_create(&cfg, &h) != ZK_OK) return fail(nullptr, "zk_create");
    if (o.slip && zk_set_slip_walls(h, 1) != ZK_OK) return fail(h, "zk_set_slip_walls");
    if (zk_load(h, in.ny, in.nx, in.da, in.gamma, in.rho.data(), in.u.data(), in.v.data(), in.p.data(),
                in.bct.data()) != ZK_OK)
        return fail(h, "zk_load");
    if (o.recon != ZK_RECON_WENO5 && zk_set_reconstruction(h, o.recon) != ZK_OK)
        return fail(h, "zk_set_reconstruction");
    zk_mlp *net = nullptr;
    if (cfg.method == ZK_METHOD_DG && o.limiter == ZK_LIMITER_LEARNED && !attach_network(h, o, &net)) {
        zk_destroy(h);
        if (net) zk_mlp_destroy(net);
        return 1;
    }

    double lam = 0.0;
    if (zk_max_speed(h, &lam) != ZK_OK) return fail(h, "zk_max_speed");
    const double hh = in.da;
    double t = 0.0, dt = 0.0, compute_ms = 0.0, io_ms = 0.0;
    int step = 0;
    // the reference writes every 100 steps (src_DG/solver.cpp:267); --output-every
    // changes that, 0 keeping the last step only
    const int output_interval = o.every > 0 ? o.every : (1 << 30);

#ifdef ZK_CLI_WENO
    std::cout << "================================\n2D Euler WENO5 solver (MI355X)\n"
              << "mesh: " << mesh_dir << "  mode: " << dt_arg << "  arg: " << third << "\n"
              << "grid: " << in.ny << " x " << in.nx << "\n================================\n";
    auto auto_rule = [&](double l) { return o.cfl > 0 ? o.cfl * hh / l : l * hh; };
    dt = (auto_dt || time_mode) ? auto_rule(lam) : dt_user;
    const int progress = std::max(1, timesteps / 1000);
    while (true) {
        if (time_mode && t >= t_end) break;
        if (!time_mode && step >= timesteps) break;
        if (auto_dt || time_mode) {
            if (zk_max_speed(h, &lam) != ZK_OK) return fail(h, "zk_max_speed");
            dt = auto_rule(lam);
            if (time_mode && t + dt > t_end) dt = t_end - t;
        }
        // fixed dt: advance to the next progress/output boundary in one call
        int chunk = 1;
        if (!auto_dt && !time_mode) {
            int next = std::min(timesteps, ((step / output_interval) + 1) * output_interval);
            if (!o.quiet) next = std::min(next, ((step / progress) + 1) * progress);
            chunk = std::max(1, next - step);
        }
        const auto c0 = Clock::now();
        if (zk_step(h, dt, chunk) != ZK_OK || zk_sync(h) != ZK_OK) return fail(h, "zk_step");
        compute_ms += ms_since(c0);
        for (int c = 0; c < chunk; ++c) t += dt;
        step += chunk;
        if (!o.quiet && step % progress == 0) {
            const double used = ms_since(wall0);
            const double eta = time_mode ? (t_end - t) / dt * used / step : used / step * (timesteps - step);
            std::printf("Step %d  t=%.6f  elapsed: %.2f s  remaining: %.2f s\n", step, t, used / 1000.0, eta / 1000.0);
        }
        if (step % output_interval == 0 || step == timesteps || (time_mode && t >= t_end)) {
            const auto i0 = Clock::now();
            save(h, in, nm, step, o.dump, o.ranks);
            io_ms += ms_since(i0);
        }
    }
#else
    const int P = ZK_CLI_DG_P;
    const double cfl_limit = 0.9 / (2.0 * (2 * P + 1));
    auto safe_dt = [&](double l) { return cfl_limit * hh / std::max(l, 1e-7); };
    std::cout << "grid: " << in.ny << " x " << in.nx << "  gamma: " << in.gamma << "  DG P = " << P
              << "  modes = " << nm << " per cell\n";
    dt = (auto_dt || time_mode) ? safe_dt(lam) : dt_user;
    const double cfl_actual = lam * dt / hh;
    std::cout << "dt = " << dt << "  h = " << hh << "  lambda_max = " << lam << "\n"
              << "DG CFL (P=" << P << ", SSP-RK3): nu_lim = " << cfl_limit << "  safe dt = " << safe_dt(lam)
              << "  nu = " << cfl_actual << "\n";
    if (cfl_actual > cfl_limit) {
        std::cerr << "[error] CFL " << cfl_actual << " exceeds the DG P=" << P << " limit " << cfl_limit
                  << "; use dt <= " << safe_dt(lam) << " or 'auto'\n";
        zk_destroy(h);
        if (net) zk_mlp_destroy(net);
        return 1;
    }
    const int print_interval = std::max(1, time_mode ? 10 : timesteps / 10000);
    while ((!time_mode && step < timesteps) || (time_mode && t < t_end)) {
        if (step > 0 && step % 10 == 0) {
            if (zk_max_speed(h, &lam) != ZK_OK) return fail(h, "zk_max_speed");
            if (auto_dt || time_mode) {
                dt = safe_dt(lam);
                if (time_mode && t + dt > t_end) dt = t_end - t;
                dt = std::max(dt, 1e-6);
            } else if (lam * dt / hh > cfl_limit) {
                std::fprintf(stderr, "[warning] step %d: CFL = %.6f exceeds %.6f (lambda_max=%.6f, dt=%.6f)\n", step,
                             lam * dt / hh, cfl_limit, lam, dt);
            }
        }
        // advance to the next recheck/progress/output boundary with one call
        int next = ((step / 10) + 1) * 10;
        next = std::min(next, ((step + 1 + output_interval - 1) / output_interval) * output_interval);
        if (!o.quiet) next = std::min(next, ((step / print_interval) + 1) * print_interval);
        if (!time_mode) next = std::min(next, timesteps);
        int chunk = std::max(1, next - step);
        if (time_mode) {
            // t must not be advanced past t_end inside a chunk
            int room = 0;
            double tt = t;
            while (room < chunk && tt < t_end) {
                tt += dt;
                ++room;
            }
            chunk = std::max(1, room);
        }
        const auto c0 = Clock::now();
        if (zk_step(h, dt, chunk) != ZK_OK || zk_sync(h) != ZK_OK) return fail(h, "zk_step");
        compute_ms += ms_since(c0);
        for (int c = 0; c < chunk; ++c) {
            t += dt;
            if (!o.quiet && step % print_interval == 0) {
                const double used = ms_since(wall0);
                const double eta = time_mode ? used / t * (t_end - t) : used / (step + 1) * ((double)timesteps - step - 1);
                std::printf("%6d / %s  dt=%.12f  t=%.6f  elapsed: %.2f s  remaining: %.2f s\n", step + 1,
                            time_mode ? "-" : std::to_string(timesteps).c_str(), dt, t, used / 1000.0, eta / 1000.0);
            }
            ++step;
        }
        if (step % output_interval == 0 || step == timesteps || (time_mode && t >= t_end)) {
            const auto i0 = Clock::now();
            save(h, in, nm, step, o.dump, o.ranks);
            io_ms += ms_since(i0);
        }
    }
#endif
    int64_t nupd = 0;
    zk_updated_cells(h, &nupd);
    const double wall_ms = ms_since(wall0);
    perf_report(wall_ms, compute_ms, io_ms, step, (long long)in.ny * in.nx,
                (time_mode ? "t_end=" + std::to_string(t_end) : "steps=" + std::to_string(timesteps)) +
                    "  t=" + std::to_string(t));
    if (compute_ms > 0)
        std::printf("cell-stage updates/s: %.4e (%lld updated cells x 3 stages x %d steps)\n",
                    3.0 * (double)nupd * step / (compute_ms / 1e3), (long long)nupd, step);
    if (net) {
        int64_t st[8];
        zk_learned_stats(h, st, 0);
        std::printf("learned indicator: %lld candidate cell-stages, %lld audited, %lld agree, %lld missed, %lld false"
                    " flags, %lld flagged by the network, %lld deferred (in band)\n", (long long)st[6],
                    (long long)st[0], (long long)st[1], (long long)st[2], (long long)st[3], (long long)st[4],
                    (long long)st[5]);
    }
    zk_destroy(h);
    if (net) zk_mlp_destroy(net);
    return 0;
}
