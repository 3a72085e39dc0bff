// Host driver with the reference's command line (init.cu:7-96) and time loop (main.cu:13-816):
// grid metrics (grid.cu:7-94), initial conditions (init.cu:99-378), adaptive dt and output-time
// clipping (main.cu:434-551), diagnostics rows (main.cu:325-421, 565-692), time-averaged profiles
// (main.cu:694-774, io.cu:262-312) and snapshots. The per-step work runs through the C-ABI:
// nse2d_channel_step (fused stage kernels + Poisson + Thom) and the diag launchers.
//
// Differences from the reference, by design:
//   - snapshots are a raw float64 .bin plus a .json header (HDF5 is not part of this stack);
//   - all diagnostic scalars of a row land in one device array and come back in one copy;
//   - the closure_diag wall bookkeeping is accumulated on the device (no per-stage host sync);
//   - the t = 0 snapshot writes u, v computed from psi (the reference writes unset buffers).
#include "../../include/nse2d_hip.h"
#include "nse2d_common.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <random>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace nse2d {
void set_last_error(const std::string& msg);
}

namespace {

using nse2d::kPi;

struct Grid {  // common.cuh:42-57
    int Nx{1024}, Ny{1025};
    double Lx{12.566370612}, h{1.0};
    std::vector<double> y, a_node, a_edge, dy_edge, w_node;
    double dx() const { return Lx / (double)Nx; }
    double dy() const { return (2.0 * h) / (double)(Ny - 1); }
    double Ly() const { return 2.0 * h; }
};

struct Params {  // common.cuh:58-114
    double Ub{1.0};
    std::string stretch{"none"}, ytable{"none"};
    double beta{0.0}, Reb{10000.0}, nu_override{-1};
    double nu(double h) const { return (nu_override >= 0 ? nu_override : (2.0 * h * Ub) / Reb); }
    double dt_init{1e-3}, dt_max{1e-2}, cfl{0.5}, cvisc{0.5};
    bool adapt{true};
    double t_end{200.0}, F0{0.0};
    int nforce{1};
    double lin_drag{0.0};
    bool resume{false};
    std::string init{"load"};
    double amp{0.5}, alpha{1.2};
    int my{1};
    double phase{0.0}, rand_amp{0.0};
    unsigned long long seed{123456789ULL};
    std::string load_path, snap_fields{"all"};
    bool load_use_psi{false};
    int rand_nx{48}, rand_mymax{8};
    double rand_k0{0.0}, rand_sigma{0.6}, rand_amp_abs{0.0};
    bool mms{false};
    double mms_A0{1e-3}, mms_eps{0.1}, mms_Om{1.0};
    int mms_kx{1};
    bool closure_diag{false};
    // this project's options
    bool free_slip{false};
    long long max_steps{-1};
    int band_kmax{32};
};

struct Run {
    Grid G;
    Params P;
    int statsEvery = 50, timeEvery = 50, spectraEvery = 0, profilesWriteEvery = 0;
    bool profile = false;
    double snap_save_dt = 0.1, diag_save_dt = 0.2;
    std::string snap_fmt = "h5", outdir = "out";
};

void parse_args(int argc, char** argv, Run& R) {  // init.cu:7-96
    Grid& G = R.G;
    Params& P = R.P;
    for (int i = 0; i < argc; i++) {
        auto eq = [&](const char* a, const char* b) { return std::strcmp(a, b) == 0; };
        const bool nx = i + 1 < argc;
        if (eq(argv[i], "--Nx") && nx) G.Nx = std::atoi(argv[++i]);
        else if (eq(argv[i], "--Ny") && nx) G.Ny = std::atoi(argv[++i]);
        else if (eq(argv[i], "--Lx") && nx) G.Lx = std::atof(argv[++i]);
        else if (eq(argv[i], "--h") && nx) G.h = std::atof(argv[++i]);
        else if (eq(argv[i], "--Ub") && nx) P.Ub = std::atof(argv[++i]);
        else if (eq(argv[i], "--Reb") && nx) P.Reb = std::atof(argv[++i]);
        else if (eq(argv[i], "--dt") && nx) P.dt_init = std::atof(argv[++i]);
        else if (eq(argv[i], "--dtmax") && nx) P.dt_max = std::atof(argv[++i]);
        else if (eq(argv[i], "--cfl") && nx) P.cfl = std::atof(argv[++i]);
        else if (eq(argv[i], "--cvisc") && nx) P.cvisc = std::atof(argv[++i]);
        else if (eq(argv[i], "--no-adapt")) P.adapt = false;
        else if (eq(argv[i], "--tend") && nx) P.t_end = std::atof(argv[++i]);
        else if (eq(argv[i], "--F0") && nx) P.F0 = std::atof(argv[++i]);
        else if (eq(argv[i], "--nforce") && nx) P.nforce = std::atoi(argv[++i]);
        else if (eq(argv[i], "--drag") && nx) P.lin_drag = std::atof(argv[++i]);
        else if (eq(argv[i], "--nu") && nx) P.nu_override = std::atof(argv[++i]);
        else if (eq(argv[i], "--stretch") && nx) P.stretch = argv[++i];
        else if (eq(argv[i], "--beta") && nx) P.beta = std::atof(argv[++i]);
        else if (eq(argv[i], "--ytable") && nx) P.ytable = argv[++i];
        else if (eq(argv[i], "--integrator") && nx) ++i;  // parsed and ignored, as in the reference
        else if (eq(argv[i], "--init") && nx) P.init = argv[++i];
        else if (eq(argv[i], "--amp") && nx) P.amp = std::atof(argv[++i]);
        else if (eq(argv[i], "--alpha") && nx) P.alpha = std::atof(argv[++i]);
        else if (eq(argv[i], "--my") && nx) P.my = std::atoi(argv[++i]);
        else if (eq(argv[i], "--phase") && nx) P.phase = std::atof(argv[++i]);
        else if (eq(argv[i], "--rand_amp") && nx) P.rand_amp = std::atof(argv[++i]);
        else if (eq(argv[i], "--seed") && nx) P.seed = (unsigned long long)std::stoull(argv[++i]);
        else if (eq(argv[i], "--load_path") && nx) P.load_path = argv[++i];
        else if (eq(argv[i], "--load_use_psi")) P.load_use_psi = true;
        else if (eq(argv[i], "--rand_nx") && nx) P.rand_nx = std::atoi(argv[++i]);
        else if (eq(argv[i], "--rand_mymax") && nx) P.rand_mymax = std::atoi(argv[++i]);
        else if (eq(argv[i], "--rand_k0") && nx) P.rand_k0 = std::atof(argv[++i]);
        else if (eq(argv[i], "--rand_sigma") && nx) P.rand_sigma = std::atof(argv[++i]);
        else if (eq(argv[i], "--rand_amp_abs") && nx) P.rand_amp_abs = std::atof(argv[++i]);
        else if (eq(argv[i], "--mms")) {
            P.mms = true;
            if (i + 1 < argc && argv[i + 1][0] != '-') P.mms = (std::atoi(argv[++i]) != 0);
        }
        else if (eq(argv[i], "--mmsA0") && nx) P.mms_A0 = std::atof(argv[++i]);
        else if (eq(argv[i], "--mmsEps") && nx) P.mms_eps = std::atof(argv[++i]);
        else if (eq(argv[i], "--mmsOm") && nx) P.mms_Om = std::atof(argv[++i]);
        else if (eq(argv[i], "--mmsKx") && nx) P.mms_kx = std::atoi(argv[++i]);
        else if (eq(argv[i], "--statsEvery") && nx) R.statsEvery = std::atoi(argv[++i]);
        else if (eq(argv[i], "--timeEvery") && nx) R.timeEvery = std::atoi(argv[++i]);
        else if (eq(argv[i], "--profile")) R.profile = true;
        else if (eq(argv[i], "--closure_diag")) P.closure_diag = true;
        else if (eq(argv[i], "--resume")) P.resume = true;
        else if (eq(argv[i], "--profilesWriteEvery") && nx) R.profilesWriteEvery = std::atoi(argv[++i]);
        else if (eq(argv[i], "--snap_save") && nx) R.snap_save_dt = std::atof(argv[++i]);
        else if (eq(argv[i], "--snap_fmt") && nx) R.snap_fmt = argv[++i];
        else if (eq(argv[i], "--snap_fields") && nx) P.snap_fields = argv[++i];
        else if (eq(argv[i], "--diag_save") && nx) R.diag_save_dt = std::atof(argv[++i]);
        else if (eq(argv[i], "--diagsave") && nx) R.diag_save_dt = std::atof(argv[++i]);
        else if (eq(argv[i], "--spectraEvery") && nx) R.spectraEvery = std::atoi(argv[++i]);
        else if (eq(argv[i], "--outdir") && nx) R.outdir = argv[++i];
        else if (eq(argv[i], "--walls") && nx) P.free_slip = (std::strcmp(argv[++i], "freeslip") == 0);
        else if (eq(argv[i], "--max_steps") && nx) P.max_steps = std::atoll(argv[++i]);
        else if (eq(argv[i], "--band_kmax") && nx) P.band_kmax = std::atoi(argv[++i]);
    }
    if (P.init == "laminar" && P.rand_amp > 0.0 && P.amp <= 0.0) { P.init = "rand"; P.amp = P.rand_amp; }
    if (P.init == "mms") P.mms = true;
}

// ---- grid.cu:7-94 (uniform, tanh, and a tabulated y read from a text file of Ny values)
bool build_y_metrics(Grid& G, const Params& P, std::string& err) {
    const int Ny = G.Ny;
    const double h = G.h;
    G.y.assign(Ny, 0.0);
    if (Ny <= 2 || P.stretch == "none" || P.stretch == "uniform") {
        const double dy = G.dy();
        for (int j = 0; j < Ny; ++j) G.y[j] = -h + j * dy;
        const double deta = 2.0 / (double)(Ny - 1);
        const double a = 1.0 / (dy / deta);
        G.a_node.assign(Ny, a);
        G.a_edge.assign(Ny - 1, a);
        G.dy_edge.assign(Ny - 1, dy);
        G.w_node.assign(Ny, dy);
        G.w_node.front() *= 0.5;
        G.w_node.back() *= 0.5;
        return true;
    }
    if (P.stretch == "tanh") {
        const double beta = (P.beta > 0 ? P.beta : 2.5);
        G.a_node.resize(Ny);
        for (int j = 0; j < Ny; ++j) {
            const double e = -1.0 + 2.0 * (double)j / (double)(Ny - 1);
            G.y[j] = h * std::tanh(beta * e) / std::tanh(beta);
            const double sech = 1.0 / std::cosh(beta * e);
            G.a_node[j] = 1.0 / (h * beta * (sech * sech) / std::tanh(beta));
        }
    } else if (P.stretch == "table") {
        std::ifstream f(P.ytable);
        if (!f) { err = "cannot open --ytable " + P.ytable; return false; }
        for (int j = 0; j < Ny; ++j)
            if (!(f >> G.y[j])) { err = "--ytable: fewer than Ny values"; return false; }
        const double deta = 2.0 / (double)(Ny - 1);
        G.a_node.resize(Ny);
        for (int j = 0; j < Ny; ++j) {
            double yeta;
            if (j == 0) yeta = (G.y[1] - G.y[0]) / deta;
            else if (j == Ny - 1) yeta = (G.y[Ny - 1] - G.y[Ny - 2]) / deta;
            else yeta = (G.y[j + 1] - G.y[j - 1]) / (2.0 * deta);
            G.a_node[j] = 1.0 / yeta;
        }
    } else {
        err = "Unknown --stretch '" + P.stretch + "'";
        return false;
    }
    G.a_edge.resize(Ny - 1);
    G.dy_edge.resize(Ny - 1);
    for (int j = 0; j < Ny - 1; ++j) {
        G.a_edge[j] = 0.5 * (G.a_node[j] + G.a_node[j + 1]);
        G.dy_edge[j] = G.y[j + 1] - G.y[j];
    }
    G.w_node.assign(Ny, 0.0);
    G.w_node[0] = 0.5 * G.dy_edge[0];
    G.w_node[Ny - 1] = 0.5 * G.dy_edge[Ny - 2];
    for (int j = 1; j < Ny - 1; ++j) G.w_node[j] = 0.5 * (G.dy_edge[j - 1] + G.dy_edge[j]);
    return true;
}

// ---- files
void ensure_dir(const std::string& p) {
    std::string cur;
    if (!p.empty() && p[0] == '/') cur = "/";
    std::stringstream ss(p);
    std::string part;
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        ::mkdir(cur.c_str(), 0755);
    }
}
bool file_nonempty(const std::string& path) {
    std::ifstream f(path, std::ios::ate | std::ios::binary);
    return f.good() && f.tellg() > 0;
}
void tee(const std::string& path, const char* fmt, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    const int n = std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (n < 0) return;
    std::fwrite(buf, 1, (size_t)std::min(n, (int)sizeof(buf) - 1), stdout);
    std::fflush(stdout);
    std::ofstream p(path, std::ios::app);
    p.write(buf, std::min(n, (int)sizeof(buf) - 1));
}

// snapshots: <outdir>/snaps/snap_t%.6f.{bin,json}; fields in order, then y
std::string snap_base(const std::string& outdir, double t) {
    std::ostringstream fn;
    fn << outdir << "/snaps/snap_t" << std::fixed << std::setprecision(6) << t;
    return fn.str();
}
void write_snapshot(const Run& R, double t, const std::vector<std::pair<std::string, const double*>>& fields) {
    const Grid& G = R.G;
    ensure_dir(R.outdir + "/snaps");
    const std::string base = snap_base(R.outdir, t);
    std::ofstream b(base + ".bin", std::ios::binary);
    const size_t N = (size_t)G.Nx * G.Ny;
    for (auto& f : fields) b.write(reinterpret_cast<const char*>(f.second), (std::streamsize)(N * sizeof(double)));
    b.write(reinterpret_cast<const char*>(G.y.data()), (std::streamsize)(G.Ny * sizeof(double)));
    std::ofstream j(base + ".json");
    j << std::setprecision(17) << "{\"t\": " << t << ", \"Nx\": " << G.Nx << ", \"Ny\": " << G.Ny
      << ", \"Lx\": " << G.Lx << ", \"Ly\": " << G.Ly() << ", \"h\": " << G.h << ", \"dx\": " << G.dx()
      << ", \"dy\": " << G.dy() << ", \"nu\": " << R.P.nu(G.h) << ", \"beta\": " << R.P.beta
      << ", \"stretch_code\": " << (R.P.stretch == "tanh" ? 1 : 0) << ", \"fields\": [";
    for (size_t k = 0; k < fields.size(); ++k) j << (k ? ", " : "") << "\"" << fields[k].first << "\"";
    j << (fields.empty() ? "" : ", ") << "\"y\"], \"dtype\": \"<f8\"}\n";
}
bool read_snapshot(const std::string& path, const std::string& name, int Nx, int Ny, double* out, double* t_out) {
    std::string base = path;
    for (const char* ext : {".json", ".bin", ".h5"})
        if (base.size() > std::strlen(ext) && base.substr(base.size() - std::strlen(ext)) == ext)
            base = base.substr(0, base.size() - std::strlen(ext));
    std::ifstream j(base + ".json");
    if (!j) return false;
    std::string s((std::istreambuf_iterator<char>(j)), std::istreambuf_iterator<char>());
    auto num = [&](const char* key) {
        const size_t p = s.find(std::string("\"") + key + "\"");
        return p == std::string::npos ? std::nan("") : std::atof(s.c_str() + s.find(':', p) + 1);
    };
    if ((int)num("Nx") != Nx || (int)num("Ny") != Ny) return false;
    if (t_out) *t_out = num("t");
    const size_t fpos = s.find("\"fields\"");
    size_t pos = s.find('[', fpos);
    const size_t end = s.find(']', pos);
    int index = -1, k = 0;
    while (pos < end) {
        const size_t q0 = s.find('"', pos);
        if (q0 == std::string::npos || q0 > end) break;
        const size_t q1 = s.find('"', q0 + 1);
        if (s.substr(q0 + 1, q1 - q0 - 1) == name) index = k;
        ++k;
        pos = q1 + 1;
    }
    if (index < 0) return false;
    if (!out) return true;
    std::ifstream b(base + ".bin", std::ios::binary);
    const size_t N = (size_t)Nx * Ny;
    b.seekg((std::streamoff)((size_t)index * N * sizeof(double)));
    b.read(reinterpret_cast<char*>(out), (std::streamsize)(N * sizeof(double)));
    return (bool)b;
}

// splitmix64 -> uniform (0,1) -> Box-Muller; seeded noise of the benchmark ICs.
struct Normal {
    unsigned long long s;
    explicit Normal(unsigned long long seed) : s(seed) {}
    unsigned long long next() {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uni() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    void pair(double& a, double& b) {
        const double u1 = uni(), u2 = uni();
        const double r = std::sqrt(-2.0 * std::log(u1));
        a = r * std::cos(2.0 * kPi * u2);
        b = r * std::sin(2.0 * kPi * u2);
    }
};

// ---- init.cu:99-378 and the benchmark-configuration ICs
bool init_omega(const Run& R, std::vector<double>& w, std::string& err) {
    const Grid& G = R.G;
    const Params& P = R.P;
    const int Nx = G.Nx, Ny = G.Ny;
    const double Lx = G.Lx, h = G.h, Ub = P.Ub;
    auto at = [Nx](int j, int i) { return (size_t)j * Nx + i; };
    w.assign((size_t)Nx * Ny, 0.0);
    const std::vector<double>& y = G.y;
    if (P.init == "mms") {
        const double k = 2.0 * kPi * (double)P.mms_kx / Lx, c = kPi / (2.0 * h);
        for (int j = 0; j < Ny; ++j) {
            const double theta = c * (y[j] + h);
            double s = std::sin(theta);
            s *= s;
            const double g = (k * k) * s - 2.0 * (c * c) * std::cos(2.0 * theta);
            for (int i = 0; i < Nx; ++i) w[at(j, i)] = P.mms_A0 * std::sin(k * ((double)i * (Lx / (double)Nx))) * g;
        }
        return true;
    }
    if (P.init == "diffusion") {
        const int my0 = (P.my > 0 ? P.my : 1);
        const double mu = (double)(2 * my0) * kPi / (2.0 * h);
        const double Au = (P.amp != 0.0) ? P.amp : 1.0;
        for (int j = 0; j < Ny; ++j) {
            const double wj = -Au * mu * std::cos(mu * (y[j] + h));
            for (int i = 0; i < Nx; ++i) w[at(j, i)] = wj;
        }
        return true;
    }
    if (P.init == "diffusion2d") {
        const int my = (P.my > 0 ? P.my : 1);
        const double alpha = (P.alpha > 0.0) ? P.alpha : (2.0 * kPi / Lx);
        const double beta = (double)my * kPi / (2.0 * h);
        const double lambda = alpha * alpha + beta * beta;
        const double Apsi = (P.amp != 0.0) ? P.amp : 1.0;
        for (int j = 0; j < Ny; ++j) {
            const double sj = std::sin(beta * (y[j] + h));
            for (int i = 0; i < Nx; ++i)
                w[at(j, i)] = lambda * (Apsi * std::sin(alpha * ((double)i * (Lx / (double)Nx))) * sj);
        }
        return true;
    }
    if (P.init == "load") {
        if (P.load_path.empty()) { err = "--init load needs --load_path"; return false; }
        if (!P.load_use_psi && read_snapshot(P.load_path, "omega", Nx, Ny, w.data(), nullptr)) return true;
        std::vector<double> psi((size_t)Nx * Ny);
        if (!read_snapshot(P.load_path, "psi", Nx, Ny, psi.data(), nullptr)) {
            err = "cannot read omega or psi from snapshot " + P.load_path;
            return false;
        }
        const double dx = G.dx(), dy = G.dy();
        const double inv_dx2 = 1.0 / (dx * dx), inv_dy2 = 1.0 / (dy * dy);
        for (int j = 1; j < Ny - 1; ++j)
            for (int i = 0; i < Nx; ++i) {
                const double pc = psi[at(j, i)];
                const double pxx = (psi[at(j, nse2d::wrap_x(i + 1, Nx))] - 2 * pc + psi[at(j, nse2d::wrap_x(i - 1, Nx))]) * inv_dx2;
                const double pyy = (psi[at(j + 1, i)] - 2 * pc + psi[at(j - 1, i)]) * inv_dy2;
                w[at(j, i)] = -(pxx + pyy);
            }
        for (int i = 0; i < Nx; ++i) {
            w[at(0, i)] = -2.0 * (psi[at(1, i)] - 0.0) * inv_dy2;
            w[at(Ny - 1, i)] = -2.0 * (psi[at(Ny - 2, i)] - 2.0 * h * Ub) * inv_dy2;
        }
        return true;
    }
    for (int j = 0; j < Ny; ++j) {  // laminar base, init.cu:284-288
        const double wj = 3.0 * Ub * y[j] / (h * h);
        for (int i = 0; i < Nx; ++i) w[at(j, i)] = wj;
    }
    unsigned long long seed = P.seed;  // init.cu:207-212
    auto rnd01 = [&]() {
        seed = 6364136223846793005ULL * seed + 1ULL;
        return ((seed >> 11) * 1.1102230246251565e-16);
    };
    if (P.init == "mix") {
        const double scale = std::max(1e-12, P.amp) * (Ub / h);
        for (double& v : w) v += scale * (rnd01() - 0.5);
    }
    if (P.init == "ts" || P.init == "mix") {
        const double alpha = (P.alpha > 0 ? P.alpha : (2.0 * kPi / Lx));
        const double mu = kPi * (double)P.my / (2.0 * h);
        const double Apsi = (P.amp * Ub) / std::max(1e-12, mu);
        for (int j = 0; j < Ny; ++j) {
            const double sy = std::sin(mu * (y[j] + h));
            for (int i = 0; i < Nx; ++i) {
                const double psi_p = Apsi * sy * std::cos(alpha * ((double)i * (Lx / (double)Nx)) + P.phase);
                w[at(j, i)] += (alpha * alpha + mu * mu) * psi_p;
            }
        }
    }
    if (P.init == "rand") {  // init.cu:314-362
        const double alpha_min = 2.0 * kPi / Lx, alpha_max = 0.35 * (kPi / G.dx());
        const int nxm = std::max(1, P.rand_nx), mymx = std::max(1, P.rand_mymax);
        const double k0 = (P.rand_k0 > 0 ? P.rand_k0 : (6.0 * 2.0 * kPi / Lx));
        const double sig = std::max(0.05, P.rand_sigma);
        std::mt19937_64 rng(P.seed);
        std::uniform_real_distribution<double> U01(0.0, 1.0);
        std::normal_distribution<double> N0(0.0, 1.0);
        const double A0 = std::max(1e-12, (P.rand_amp_abs > 0 ? P.rand_amp_abs : P.amp)) / std::sqrt((double)nxm);
        std::vector<std::vector<double>> sY(mymx + 1, std::vector<double>(Ny, 0.0));
        for (int my = 1; my <= mymx; ++my)
            for (int j = 0; j < Ny; ++j) sY[my][j] = std::sin(kPi * (double)my / (2.0 * h) * (y[j] + h));
        for (int m = 0; m < nxm; ++m) {
            double alpha = std::exp(std::log(k0) + sig * N0(rng));
            if (alpha < alpha_min) alpha = alpha_min;
            if (alpha > alpha_max) alpha = alpha_max;
            const int my = 1 + (int)std::floor(U01(rng) * mymx);
            const double phi = 2.0 * kPi * U01(rng);
            const double mu = kPi * (double)my / (2.0 * h);
            const double k = std::sqrt(alpha * alpha + mu * mu);
            const double wamp = A0 * std::exp(-0.5 * std::pow(std::log(k / k0) / sig, 2));
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i)
                    w[at(j, i)] += wamp * std::cos(alpha * ((double)i * (Lx / (double)Nx)) + phi) * sY[my][j];
        }
    }
    if (P.init == "lamnoise") {
        Normal g(P.seed);
        const size_t N = w.size();
        double a = 0, b = 0;
        for (size_t idx = 0; idx < N; ++idx) {
            if ((idx & 1) == 0) g.pair(a, b);
            const double yy = y[idx / (size_t)Nx] / h;
            const double taper = (1.0 - yy * yy) * (1.0 - yy * yy);
            w[idx] += P.amp * taper * ((idx & 1) ? b : a);
        }
    }
    if (P.init == "lamband") {
        Normal g(P.seed);
        const int K = std::max(1, P.band_kmax);
        std::vector<double> ct(Nx), st(Nx), ca((size_t)Ny * K), cb((size_t)Ny * K);
        for (int m = 0; m < Nx; ++m) {
            ct[m] = std::cos(2.0 * kPi * (double)m / (double)Nx);
            st[m] = std::sin(2.0 * kPi * (double)m / (double)Nx);
        }
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < K; ++k) g.pair(ca[(size_t)j * K + k], cb[(size_t)j * K + k]);
        const double norm = 1.0 / std::sqrt((double)K);
        for (int j = 0; j < Ny; ++j) {
            const double yy = y[j] / h, taper = (1.0 - yy * yy) * (1.0 - yy * yy);
            for (int i = 0; i < Nx; ++i) {
                double s = 0.0;
                for (int k = 1; k <= K; ++k) {
                    const int m = (int)(((long long)k * i) % Nx);
                    s += ca[(size_t)j * K + k - 1] * ct[m] + cb[(size_t)j * K + k - 1] * st[m];
                }
                w[at(j, i)] += P.amp * taper * (s * norm);
            }
        }
    }
    if (P.init != "laminar" && P.init != "mix" && P.init != "ts" && P.init != "rand" &&
        P.init != "lamnoise" && P.init != "lamband") {
        err = "unknown --init " + P.init;
        return false;
    }
    return true;
}

// ---- io.cu:162-178
double time_tol(double t) { return 64.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, std::fabs(t)); }
double sched_tol(double t, double dt) { return std::max(time_tol(t), 1e-6 * std::max(1e-300, std::fabs(dt))); }
double next_after(double t0, double dt) {
    if (dt <= 0.0) return 1e300;
    const double k = std::floor((t0 + time_tol(t0)) / dt);
    const double t1 = (k + 1.0) * dt;
    return (t1 <= t0 + time_tol(t0)) ? t1 + dt : t1;
}

// ---- io.cu:262-312
void write_profiles_csv(const std::string& path, const Grid& G, const Params& P, double utau_mean, double T,
                        const std::vector<double>* H) {
    std::ofstream f(path, std::ios::out);
    f << "j,y,yplus,U,V,uu,vv,uv,Uplus,tau_total\n";
    const bool mapped_out = (P.stretch != "none" && P.beta > 0.0);
    const double nu = P.nu(G.h);
    for (int j = 0; j < G.Ny; ++j) {
        const double y = mapped_out ? G.y[j] : (-G.h + (double)j * G.dy());
        const double Ub = (T > 0 ? H[0][j] / T : 0.0), Vb = (T > 0 ? H[1][j] / T : 0.0);
        const double U2 = (T > 0 ? H[2][j] / T : 0.0), V2 = (T > 0 ? H[3][j] / T : 0.0);
        const double UV = (T > 0 ? H[4][j] / T : 0.0);
        const double uu = std::fmax(0.0, U2 - Ub * Ub), vv = std::fmax(0.0, V2 - Vb * Vb), uv = UV - Ub * Vb;
        const double Ujp = (T > 0 ? H[0][std::min(j + 1, G.Ny - 1)] / T : 0.0);
        const double Ujm = (T > 0 ? H[0][std::max(j - 1, 0)] / T : 0.0);
        double dUdy;
        if (mapped_out) {
            const double inv_deta = (G.Ny > 1) ? (0.5 * (G.Ny - 1)) : 0.0;
            const double a = G.a_node[j];
            if (j == 0) dUdy = a * (Ujp - Ub) * inv_deta;
            else if (j == G.Ny - 1) dUdy = a * (Ub - Ujm) * inv_deta;
            else dUdy = a * (Ujp - Ujm) * 0.5 * inv_deta;
        } else {
            if (j == 0) dUdy = (Ujp - Ub) / G.dy();
            else if (j == G.Ny - 1) dUdy = (Ub - Ujm) / G.dy();
            else dUdy = (Ujp - Ujm) * 0.5 / G.dy();
        }
        f.setf(std::ios::scientific);
        f.precision(8);
        f << j << "," << y << "," << (utau_mean > 0 ? utau_mean * (y + G.h) / nu : 0.0) << "," << Ub << "," << Vb
          << "," << uu << "," << vv << "," << uv << "," << (utau_mean > 0 ? Ub / utau_mean : 0.0) << ","
          << (nu * dUdy - uv) << "\n";
    }
}

template <class T>
T* dalloc(size_t n) {
    return static_cast<T*>(nse2d_dev_alloc(n * sizeof(T)));
}

// Slots of the diagnostics scalar array (one D2H copy per row).
enum Slot {
    // the first NSE2D_DIAG_NOUT slots are nse2d_diag_row's outputs
    S_UMAX = NSE2D_DIAG_UMAX, S_VMAX = NSE2D_DIAG_VMAX, S_VMAXM = NSE2D_DIAG_VMAXM, S_K = NSE2D_DIAG_K,
    S_W2 = NSE2D_DIAG_OMEGA_W2, S_GW2 = NSE2D_DIAG_OMEGA_GRAD2, S_ORHS_MIXED = NSE2D_DIAG_OMEGA_RHS,
    S_PIN = NSE2D_DIAG_PIN, S_EPS = NSE2D_DIAG_EPS, S_POM = NSE2D_DIAG_POMEGA,
    S_WWYB = NSE2D_DIAG_WALL_OMEGA_OMEGAY_BOT, S_WWYT = NSE2D_DIAG_WALL_OMEGA_OMEGAY_TOP,
    S_MWB = NSE2D_DIAG_WALL_OMEGA_BOT, S_MWT = NSE2D_DIAG_WALL_OMEGA_TOP,
    S_ORHS = NSE2D_DIAG_NOUT, S_W2I, S_GW2I, S_T2D, S_T2J, S_T2W2, S_T2G2, S_T2P, S_T2ETA, S_T2PHI, S_NSLOTS
};

int run(Run& R) {
    Grid& G = R.G;
    Params& P = R.P;
    std::string err;
    if (!build_y_metrics(G, P, err)) { std::fprintf(stderr, "ERROR: %s\n", err.c_str()); return 1; }
    if (G.Ny < 4 || G.Nx < 4) { std::fprintf(stderr, "ERROR: grid too small\n"); return 1; }
    ensure_dir(R.outdir);
    const std::string diag_path = R.outdir + "/diagnostics.csv", prof_path = R.outdir + "/profiles_timeavg.csv";
    const std::string time_path = R.outdir + "/timing.csv", prog = R.outdir + "/progress.log";
    const double nu = P.nu(G.h);
    const int Nx = G.Nx, Ny = G.Ny;
    const size_t N = (size_t)Nx * Ny;
    const double dx = G.dx(), inv_dx = 1.0 / dx, inv_dx2 = inv_dx * inv_dx;
    const double deta = 2.0 / (double)(Ny - 1), inv_deta = 1.0 / deta, inv_deta2 = inv_deta * inv_deta;
    double dy_min = std::fabs(G.dy_edge[0]);  // main.cu:56-62
    for (int j = 1; j < Ny - 1; ++j) dy_min = std::min(dy_min, std::fabs(G.dy_edge[j]));
    const double inv_dymin2 = 1.0 / (dy_min * dy_min);
    double Ly_full = 0.0;  // main.cu:64-71
    for (int j = 0; j < Ny; ++j) Ly_full += G.w_node[j];
    const double invA_full = 1.0 / ((double)Nx * Ly_full);
    double meanFx_A = 0.0;  // main.cu:78-87
    if (P.F0 != 0.0) {
        for (int j = 0; j < Ny; ++j)
            meanFx_A += P.F0 * std::sin((double)P.nforce * kPi * ((G.y[j] + G.h) / (2.0 * G.h))) * G.w_node[j];
        meanFx_A /= Ly_full;
    }

    double* d_a_node = dalloc<double>(Ny);
    double* d_a_edge = dalloc<double>(Ny - 1);
    double* d_w_node = dalloc<double>(Ny);
    double* d_y_node = dalloc<double>(Ny);
    double* d_dy_edge = dalloc<double>(Ny - 1);
    double* d_wt_node = dalloc<double>(Ny);
    nse2d_memcpy_h2d(d_a_node, G.a_node.data(), Ny * sizeof(double));
    nse2d_memcpy_h2d(d_a_edge, G.a_edge.data(), (Ny - 1) * sizeof(double));
    nse2d_memcpy_h2d(d_w_node, G.w_node.data(), Ny * sizeof(double));
    nse2d_memcpy_h2d(d_y_node, G.y.data(), Ny * sizeof(double));
    nse2d_memcpy_h2d(d_dy_edge, G.dy_edge.data(), (Ny - 1) * sizeof(double));
    {
        std::vector<double> wt(Ny, 0.0);  // main.cu:252-254
        for (int j = 1; j < Ny - 1; ++j) wt[j] = (2.0 / (double)(Ny - 1)) / G.a_node[j];
        nse2d_memcpy_h2d(d_wt_node, wt.data(), Ny * sizeof(double));
    }
    {
        std::ofstream pl(prog, std::ios::app);
    }
    tee(prog, "[grid-y] stretch=%s beta=%.3f (mapped)\n", P.stretch.c_str(), P.beta);
    tee(prog, "Z2Z channel DNS | Nx=%d Ny=%d Lx=%.6f h=%.6f\n", Nx, Ny, G.Lx, G.h);
    tee(prog, "Reb=%.6g, Ub=%.6g, nu=%.6g%s\n", P.Reb, P.Ub, nu, (P.nu_override > 0 ? " (override)" : " (from 2h*Ub/Reb)"));
    tee(prog, "[hip] %s, walls=%s\n", nse2d_version(), P.free_slip ? "free-slip" : "no-slip");

    hipStream_t stream = nullptr;
    NSE2D_HIP_CHECK(hipStreamCreate(&stream));
    nse2d_channel_desc cd{};
    cd.Nx = Nx; cd.Ny = Ny; cd.Lx = G.Lx; cd.h = G.h;
    cd.nu = nu; cd.F0 = P.F0; cd.lin_drag = P.lin_drag; cd.nforce = P.nforce;
    cd.psi_bot = 0.0; cd.psi_top = 2.0 * G.h * P.Ub;
    cd.mms = P.mms ? 1 : 0; cd.mms_A0 = P.mms_A0; cd.mms_eps = P.mms_eps; cd.mms_Om = P.mms_Om; cd.mms_kx = P.mms_kx;
    cd.cfl = P.cfl; cd.cvisc = P.cvisc; cd.dt_max = P.dt_max; cd.dy_min = dy_min;
    cd.d_a_node = d_a_node; cd.d_a_edge = d_a_edge; cd.d_y_node = d_y_node; cd.d_dy_edge = d_dy_edge; cd.d_w_node = d_w_node;
    cd.wall_type = P.free_slip ? 2 : 1;  // this run's closure; the library-wide default is left alone
    nse2d_channel_t ch = nse2d_channel_create(&cd, stream);
    if (!ch) { std::fprintf(stderr, "ERROR: %s\n", nse2d_last_error()); return 1; }
    double* d_w = nse2d_channel_field(ch, NSE2D_FIELD_OMEGA);
    double* d_psi = nse2d_channel_field(ch, NSE2D_FIELD_PSI);
    double* d_rhs = nse2d_channel_field(ch, NSE2D_FIELD_RHS);
    double* d_J = dalloc<double>(N);
    double* d_lap = dalloc<double>(N);
    double* d_u = dalloc<double>(N);
    double* d_v = dalloc<double>(N);
    double* d_sc = dalloc<double>(S_NSLOTS);
    double* d_mom = dalloc<double>(5 * (size_t)Ny);
    // x-spectra machinery (main.cu:265-268)
    nse2d_spectra1d_t spec = nullptr;
    if (R.spectraEvery > 0) {
        spec = nse2d_spectra1d_create(Nx, Ny, G.Lx, stream);
        if (!spec) { std::fprintf(stderr, "ERROR: %s\n", nse2d_last_error()); return 1; }
    }

    double t_start = 0.0;
    {
        std::vector<double> h_w;
        if (!init_omega(R, h_w, err)) { std::fprintf(stderr, "ERROR: %s\n", err.c_str()); return 1; }
        if (P.resume) {
            if (!read_snapshot(P.load_path, "omega", Nx, Ny, nullptr, &t_start) &&
                !read_snapshot(P.load_path, "psi", Nx, Ny, nullptr, &t_start)) {
                std::fprintf(stderr, "ERROR: --resume needs a readable --load_path snapshot\n");
                return 1;
            }
        }
        nse2d_memcpy_h2d(d_w, h_w.data(), N * sizeof(double));
    }
    if (!file_nonempty(diag_path)) {
        std::ofstream f(diag_path, std::ios::out);
        f << "step,t,dt,umax,vmax,K,Omega,dOmega_rhs,eta,BOmega_wall,Pin,epsilon,POmega,tauw_bot,tauw_top,dpdx,"
             "utau_avg,Cf,Re_tau";
        if (P.closure_diag)
            f << ",dOmega_rhs_same,dOmega_rhs_same_int,Omega_int,eta_int,Jwall_s1_cum,Jwall_s2_cum,Jwall_s3_cum,"
                 "Omega_int_w,dOmega_same_int_w,A_w,P_int_w,eta_h,Phi_h";
        f << "\n";
    }
    if (R.profile && !file_nonempty(time_path)) {
        std::ofstream g(time_path, std::ios::out);
        g << "step,t,dt,t_step_ms,t_fft_ms,t_tridiag_ms\n";
    }
    { std::ofstream f(prof_path, std::ios::out); f << "j,y,yplus,U,V,uu,vv,uv,Uplus,tau_total\n"; }

    nse2d_channel_init(ch);  // main.cu:211-213

    auto rebuild_rhs = [&](double t_rhs) {  // J, Laplacian, RHS (+ MMS) at the current state
        launch_arakawa_J_mapped(d_psi, d_w, d_J, Nx, Ny, inv_dx, inv_deta, d_a_node, stream);
        launch_laplacian_fd_mapped(d_w, d_lap, Nx, Ny, inv_dx2, inv_deta2, d_a_node, d_a_edge, stream);
        launch_rhs_mapped(d_J, d_lap, d_w, d_rhs, Nx, Ny, nu, P.F0, P.nforce, G.h, P.lin_drag, d_y_node, stream);
        if (P.mms)
            launch_mms_add_forcing(d_rhs, Nx, Ny, G.Lx, G.h, t_rhs, nu, P.lin_drag, P.mms_A0, P.mms_eps, P.mms_Om,
                                   P.mms_kx, d_y_node, stream);
    };
    std::vector<double> h_a(N), h_b(N), h_c(N), h_d(N);
    auto snapshot = [&](double t) {
        std::vector<std::pair<std::string, const double*>> fs;
        NSE2D_HIP_CHECK(hipStreamSynchronize(stream));
        nse2d_memcpy_d2h(h_a.data(), d_w, N * sizeof(double));
        fs.push_back({"omega", h_a.data()});
        if (P.snap_fields == "all" || P.snap_fields == "omega_psi") {
            nse2d_memcpy_d2h(h_b.data(), d_psi, N * sizeof(double));
            fs.push_back({"psi", h_b.data()});
        }
        if (P.snap_fields == "all") {
            launch_uv_from_psi_mapped(d_psi, d_u, d_v, Nx, Ny, inv_dx, inv_deta, d_a_node, stream);
            NSE2D_HIP_CHECK(hipStreamSynchronize(stream));
            nse2d_memcpy_d2h(h_c.data(), d_u, N * sizeof(double));
            nse2d_memcpy_d2h(h_d.data(), d_v, N * sizeof(double));
            fs.push_back({"u", h_c.data()});
            fs.push_back({"v", h_d.data()});
        }
        write_snapshot(R, t, fs);
    };

    const bool use_snap = (R.snap_save_dt > 0), use_diag = (R.diag_save_dt > 0);
    double next_snap_t = use_snap ? (P.resume ? next_after(t_start, R.snap_save_dt) : 0.0) : 1e300;
    double next_diag_t = use_diag ? (P.resume ? next_after(t_start, R.diag_save_dt) : 0.0) : 1e300;
    double tauw_abs_dt_sum = 0.0, T_diag = 0.0, last_diag_t = t_start;
    double* sc = nullptr;
    NSE2D_HIP_CHECK(hipHostMalloc((void**)&sc, S_NSLOTS * sizeof(double), hipHostMallocDefault));

    // One diagnostics row (main.cu:325-421 for t = 0, main.cu:565-692 afterwards).
    auto diag_row = [&](long long stepno, double t, double dt_col, bool t0) {
        if (t0) rebuild_rhs(t_start);  // main.cu:345-362: the t = 0 row builds the RHS
        // CFL maxima, K, Omega, eta, <omega rhs> (with the stored RHS), Pin, epsilon, P_Omega and the
        // wall means in one pass (the launchers of main.cu:565-692, fused; equal bit for bit)
        if (nse2d_diag_row(d_psi, d_w, d_rhs, Nx, Ny, inv_dx, inv_deta, G.h, P.F0, P.nforce, d_a_node, d_a_edge,
                           d_y_node, d_w_node, d_sc, nullptr, stream) != 0) {
            std::fprintf(stderr, "ERROR: %s\n", nse2d_last_error());
            std::abort();
        }
        if (!t0 && P.closure_diag) {  // main.cu:591-611
            rebuild_rhs(t);
            launch_omega_rhs_balance_mapped(d_w, d_rhs, Nx, Ny, d_w_node, d_sc + S_ORHS, stream);
            launch_enstrophy_diss_mapped(d_w, Nx, Ny, inv_dx, inv_deta, d_a_node, d_w_node, d_sc + S_W2I,
                                         d_sc + S_GW2I, stream);
            launch_omega_rhs_balance_mapped(d_w, d_rhs, Nx, Ny, d_wt_node, d_sc + S_T2D, stream);
            launch_omega_rhs_balance_mapped(d_w, d_J, Nx, Ny, d_wt_node, d_sc + S_T2J, stream);
            launch_enstrophy_diss_mapped(d_w, Nx, Ny, inv_dx, inv_deta, d_a_node, d_wt_node, d_sc + S_T2W2,
                                         d_sc + S_T2G2, stream);
            launch_enstrophy_input_mapped_full(d_w, Nx, Ny, G.h, P.F0, P.nforce, d_y_node, d_wt_node, d_sc + S_T2P,
                                               stream);
            launch_enstrophy_sbp_mapped(d_w, Nx, Ny, inv_dx, inv_deta, d_a_edge, d_wt_node, d_sc + S_T2ETA,
                                        d_sc + S_T2PHI, stream);
        }
        NSE2D_HIP_CHECK(hipMemcpyAsync(sc, d_sc, S_NSLOTS * sizeof(double), hipMemcpyDeviceToHost, stream));
        NSE2D_HIP_CHECK(hipStreamSynchronize(stream));
        // wall-row part of <omega rhs>_A at the same state (main.cu:290-296, 601)
        double wall_omega_rhs_same = 0.0;
        if (!t0 && P.closure_diag) {
            std::vector<double> rb(Nx), rt(Nx), wb(Nx), wtp(Nx);
            nse2d_memcpy_d2h(wb.data(), d_w, Nx * sizeof(double));
            nse2d_memcpy_d2h(wtp.data(), d_w + (size_t)(Ny - 1) * Nx, Nx * sizeof(double));
            nse2d_memcpy_d2h(rb.data(), d_rhs, Nx * sizeof(double));
            nse2d_memcpy_d2h(rt.data(), d_rhs + (size_t)(Ny - 1) * Nx, Nx * sizeof(double));
            double s = 0.0;
            for (int i = 0; i < Nx; ++i) s += G.w_node[0] * wb[i] * rb[i] + G.w_node[Ny - 1] * wtp[i] * rt[i];
            wall_omega_rhs_same = s * invA_full;
        }
        const double qnan = std::numeric_limits<double>::quiet_NaN();
        const double K = sc[S_K] * invA_full, Omega = 0.5 * sc[S_W2] * invA_full, eta = nu * sc[S_GW2] * invA_full;
        const double BOmega_wall = (nu > 0.0) ? (nu * (sc[S_WWYT] - sc[S_WWYB]) / (2.0 * G.h)) : 0.0;
        const double dOmega_rhs = sc[S_ORHS_MIXED] * invA_full;
        const double Pin = sc[S_PIN] * invA_full, epsilon = nu * sc[S_EPS] * invA_full, P_Omega = sc[S_POM] * invA_full;
        const double tauw_bot = -nu * sc[S_MWB], tauw_top = -nu * sc[S_MWT];
        const double utau = std::sqrt(0.5 * (std::fabs(tauw_bot) + std::fabs(tauw_top)));
        const double Cf = (P.Ub != 0.0) ? 2.0 * (0.5 * (std::fabs(tauw_bot) + std::fabs(tauw_top))) / (P.Ub * P.Ub) : qnan;
        const double Re_tau = (nu > 0.0) ? utau * G.h / nu : 0.0;
        const double dpdx = (tauw_top - tauw_bot) / (2.0 * G.h) + meanFx_A - P.lin_drag * P.Ub;
        if (!t0) {
            double dtw = next_diag_t - last_diag_t;
            if (dtw < 0) dtw = 0;
            last_diag_t = next_diag_t;
            if (dtw > 0) { tauw_abs_dt_sum += 0.5 * (std::fabs(tauw_bot) + std::fabs(tauw_top)) * dtw; T_diag += dtw; }
        }
        std::ofstream f(diag_path, std::ios::app);
        f.setf(std::ios::scientific);
        f.precision(8);
        f << stepno << "," << t << "," << dt_col << "," << sc[S_UMAX] << "," << sc[S_VMAX] << "," << K << "," << Omega
          << "," << dOmega_rhs << "," << eta << "," << BOmega_wall << "," << Pin << "," << epsilon << "," << P_Omega
          << "," << tauw_bot << "," << tauw_top << "," << dpdx << "," << utau << "," << Cf << "," << Re_tau;
        if (P.closure_diag) {
            if (t0) {
                f << "," << dOmega_rhs << ",nan,nan,nan," << 0.0 << "," << 0.0 << "," << 0.0 << ",nan,nan,nan,nan,nan,nan";
            } else {
                double jw[3];
                nse2d_channel_wall_bookkeeping(ch, jw, invA_full);
                const double same = sc[S_ORHS] * invA_full;
                f << "," << same << "," << (same - wall_omega_rhs_same) << "," << 0.5 * sc[S_W2I] * invA_full << ","
                  << nu * sc[S_GW2I] * invA_full << "," << jw[0] << "," << jw[1] << "," << jw[2] << ","
                  << 0.5 * sc[S_T2W2] * invA_full << "," << sc[S_T2D] * invA_full << "," << -sc[S_T2J] * invA_full
                  << "," << sc[S_T2P] * invA_full << "," << nu * sc[S_T2ETA] * invA_full << ","
                  << nu * sc[S_T2PHI] * invA_full;
            }
        }
        f << "\n";
    };

    if (!P.resume && use_snap) { snapshot(0.0); next_snap_t = R.snap_save_dt; }
    if (!P.resume && use_diag) { diag_row(0, 0.0, P.dt_init, true); next_diag_t = R.diag_save_dt; }
    tee(prog, "[run] t0=%.6f -> t_end=%.6f\n", t_start, P.t_end);

    std::vector<double> H[5];
    for (auto& v : H) v.assign(Ny, 0.0);
    std::vector<double> mom(5 * (size_t)Ny);
    double T_accum = 0.0, last_stats_t = t_start;
    double t = t_start, dt = P.dt_init;
    long long step = 0;
    hipEvent_t e0, e1;
    NSE2D_HIP_CHECK(hipEventCreate(&e0));
    NSE2D_HIP_CHECK(hipEventCreate(&e1));
    const auto wall0 = std::chrono::steady_clock::now();
    while (t < P.t_end - 0.5 * dt) {  // main.cu:434
        if (P.max_steps >= 0 && step >= P.max_steps) break;
        if (P.adapt) {  // main.cu:440-458, maxima read back once per step as in the reference
            launch_max_uv_from_psi_mapped(d_psi, Nx, Ny, inv_dx, inv_deta, d_a_node, d_sc + S_UMAX, d_sc + S_VMAX,
                                          d_sc + S_VMAXM, stream);
            NSE2D_HIP_CHECK(hipMemcpyAsync(sc, d_sc, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
            NSE2D_HIP_CHECK(hipStreamSynchronize(stream));
            const double umax = sc[S_UMAX], vmaxm = sc[S_VMAXM];
            const double dt_x = (umax > 0 ? P.cfl * (dx / umax) : 1e30);
            const double dt_y = (vmaxm > 0 ? P.cfl * (deta / vmaxm) : 1e30);
            const double dt_vis = (nu > 0) ? P.cvisc * 0.5 / (nu * (inv_dx2 + inv_dymin2)) : 1e30;
            dt = std::fmin(P.dt_max, std::fmin(std::fmin(dt_x, dt_y), dt_vis));
            if (step % 5000 == 0)
                tee(prog, "[dt] step=%lld t=%.6g umax=%.3e vmaxm=%.3e dt=%.3e\n", step, t, umax, vmaxm, dt);
        }
        if (P.adapt && (use_snap || use_diag)) {  // main.cu:467-480
            double t_next = 1e300;
            if (use_snap) t_next = std::min(t_next, next_snap_t);
            if (use_diag) t_next = std::min(t_next, next_diag_t);
            if (t_next < 1e299) {
                const double rem = t_next - t;
                const double tol = std::max(time_tol(t_next), 1e-6 * dt);
                if (rem > tol && rem < 0.5 * dt) dt = rem;
            }
        }
        // time after this step, with main.cu:544-551's snapping, to know whether a diag row follows
        double t_after = t + dt;
        if (use_snap && std::fabs(t_after - next_snap_t) <= sched_tol(t_after, dt)) t_after = next_snap_t;
        if (use_diag && std::fabs(t_after - next_diag_t) <= sched_tol(t_after, dt)) t_after = next_diag_t;
        const bool diag_next = use_diag && (t_after + time_tol(t_after) >= next_diag_t);
        int flags = 0;
        if (diag_next) flags |= NSE2D_STEP_STORE_RHS;
        if (P.closure_diag) flags |= NSE2D_STEP_WALL_BOOKKEEPING;
        if (R.profile) NSE2D_HIP_CHECK(hipEventRecord(e0, stream));
        nse2d_channel_step(ch, t, dt, flags);
        if (R.profile) { NSE2D_HIP_CHECK(hipEventRecord(e1, stream)); NSE2D_HIP_CHECK(hipEventSynchronize(e1)); }
        t = t_after;
        step++;
        if (use_snap) {
            while (t + sched_tol(t, dt) >= next_snap_t && next_snap_t <= P.t_end + sched_tol(P.t_end, dt)) {
                snapshot(next_snap_t);
                next_snap_t += R.snap_save_dt;
            }
        }
        if (diag_next) {
            diag_row(step, t, dt, false);
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
            tee(prog, "diag t=%.6f step=%lld dt=%.6e elapsed=%.3fs\n", next_diag_t, step, dt, el);
            next_diag_t += R.diag_save_dt;
        }
        if (step % R.statsEvery == 0) {  // main.cu:695-733
            const double dtw = t - last_stats_t;
            if (dtw > 0) {
                launch_row_moments_from_psi_mapped(d_psi, Nx, Ny, inv_dx, inv_deta, d_a_node, d_mom, d_mom + Ny,
                                                   d_mom + 2 * (size_t)Ny, d_mom + 3 * (size_t)Ny,
                                                   d_mom + 4 * (size_t)Ny, stream);
                NSE2D_HIP_CHECK(hipStreamSynchronize(stream));
                nse2d_memcpy_d2h(mom.data(), d_mom, 5 * (size_t)Ny * sizeof(double));
                for (int k = 0; k < 5; ++k)
                    for (int j = 0; j < Ny; ++j) H[k][j] += mom[(size_t)k * Ny + j] / (double)Nx * dtw;
                T_accum += dtw;
                last_stats_t = t;
                if (R.profilesWriteEvery > 0 && (step % R.profilesWriteEvery) == 0) {
                    double ul;
                    if (T_diag > 0) ul = std::sqrt(tauw_abs_dt_sum / T_diag);
                    else {
                        launch_mean_wall_omega(d_w, Nx, Ny, d_sc + S_MWB, d_sc + S_MWT, stream);
                        NSE2D_HIP_CHECK(hipMemcpyAsync(sc, d_sc, S_NSLOTS * sizeof(double), hipMemcpyDeviceToHost, stream));
                        NSE2D_HIP_CHECK(hipStreamSynchronize(stream));
                        ul = std::sqrt(0.5 * (std::fabs(-nu * sc[S_MWB]) + std::fabs(-nu * sc[S_MWT])));
                    }
                    write_profiles_csv(prof_path, G, P, ul, T_accum, H);
                }
            }
        }
        if (spec && (step % R.spectraEvery == 0)) {  // main.cu:735-744
            launch_uv_from_psi_mapped(d_psi, d_u, d_v, Nx, Ny, inv_dx, inv_deta, d_a_node, stream);
            nse2d_spectra1d_accumulate(spec, d_u, d_v, stream);
        }
        if (R.profile && (step % R.timeEvery == 0)) {
            float ms = 0.f;
            NSE2D_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            std::ofstream g(time_path, std::ios::app);
            g.setf(std::ios::scientific);
            g.precision(6);
            g << step << "," << t << "," << dt << "," << ms << "," << 0.0 << "," << 0.0 << "\n";
        }
    }
    NSE2D_HIP_CHECK(hipStreamSynchronize(stream));
    double utau_mean;  // main.cu:757-774
    if (T_diag > 0) utau_mean = std::sqrt(tauw_abs_dt_sum / T_diag);
    else {
        launch_mean_wall_omega(d_w, Nx, Ny, d_sc + S_MWB, d_sc + S_MWT, stream);
        NSE2D_HIP_CHECK(hipMemcpyAsync(sc, d_sc, S_NSLOTS * sizeof(double), hipMemcpyDeviceToHost, stream));
        NSE2D_HIP_CHECK(hipStreamSynchronize(stream));
        utau_mean = std::sqrt(0.5 * (std::fabs(-nu * sc[S_MWB]) + std::fabs(-nu * sc[S_MWT])));
    }
    write_profiles_csv(prof_path, G, P, utau_mean, T_accum, H);
    if (spec) {  // main.cu:776-780, diag.cuh:175-190
        const int nk = Nx / 2 + 1;
        std::vector<double> kx(nk), eu(nk), ev(nk);
        nse2d_spectra1d_read(spec, kx.data(), eu.data(), ev.data(), stream);
        std::ofstream f(R.outdir + "/spectra_kx.csv");
        f.setf(std::ios::scientific);
        f.precision(10);
        f << "kx,E_u,E_v,E_tot\n";
        for (int k = 0; k < nk; ++k) f << kx[k] << "," << eu[k] << "," << ev[k] << "," << (eu[k] + ev[k]) << "\n";
        nse2d_spectra1d_destroy(spec);
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    tee(prog, "Done. steps=%lld t=%.6f wall=%.3fs\n", step, t, el);

    NSE2D_HIP_CHECK(hipEventDestroy(e0));
    NSE2D_HIP_CHECK(hipEventDestroy(e1));
    NSE2D_HIP_CHECK(hipHostFree(sc));
    nse2d_channel_destroy(ch);
    for (double* p : {d_a_node, d_a_edge, d_w_node, d_y_node, d_dy_edge, d_wt_node, d_J, d_lap, d_u, d_v, d_sc, d_mom})
        nse2d_dev_free(p);
    NSE2D_HIP_CHECK(hipStreamDestroy(stream));
    return 0;
}

}  // namespace

extern "C" int nse2d_main(int argc, char** argv) {
    Run R;
    parse_args(argc, argv, R);
    try {
        return run(R);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        return 1;
    }
}

extern "C" int nse2d_host_metrics(int Ny, double h, const char* stretch, double beta, double* y, double* a_node,
                                  double* a_edge, double* dy_edge, double* w_node) {
    Grid G;
    G.Ny = Ny;
    G.h = h;
    Params P;
    P.stretch = stretch ? stretch : "none";
    if (P.stretch.compare(0, 6, "table:") == 0) {  // "table:<file>": --stretch table --ytable <file>
        P.ytable = P.stretch.substr(6);
        P.stretch = "table";
    }
    P.beta = beta;
    std::string err;
    if (Ny < 3 || !build_y_metrics(G, P, err)) {
        nse2d::set_last_error(err.empty() ? "nse2d_host_metrics: Ny < 3" : err);
        return 1;
    }
    std::copy(G.y.begin(), G.y.end(), y);
    std::copy(G.a_node.begin(), G.a_node.end(), a_node);
    std::copy(G.a_edge.begin(), G.a_edge.end(), a_edge);
    std::copy(G.dy_edge.begin(), G.dy_edge.end(), dy_edge);
    std::copy(G.w_node.begin(), G.w_node.end(), w_node);
    return 0;
}

extern "C" int nse2d_host_init_omega(int argc, char** argv, double* w_out) {
    Run R;
    parse_args(argc, argv, R);
    std::string err;
    std::vector<double> w;
    if (!build_y_metrics(R.G, R.P, err) || !init_omega(R, w, err)) {
        nse2d::set_last_error(err);
        return 1;
    }
    std::copy(w.begin(), w.end(), w_out);
    return 0;
}
