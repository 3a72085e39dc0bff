// The SSPRK(3,3) step of the reference driver (main.cu:482-538) on the device:
//   for s = 1..3:  stage kernel (Arakawa J + mapped Laplacian + RHS + forcing + update, fused)
//                  -> Poisson solve -> Thom wall closure
// plus the device-side adaptive dt (main.cu:440-458) so that steps need no host round trip.
#include "nse2d_common.hpp"
#include "nse2d_ops.hpp"
#include "nse2d_poisson.hpp"
#include "../../include/nse2d_hip.h"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

namespace nse2d {

void set_last_error(const std::string& msg);
void launch_max_uv(const double* psi, int Nx, int Ny, double inv_dx, double inv_deta, const double* a_node,
                   double* umax, double* vmax, double* vmaxm, hipStream_t s);

namespace {

struct StageArgs {
    const double* psi;   // streamfunction of the stage state
    const double* ws;    // stage state omega_s
    const double* wn;    // omega^n (stages 2, 3)
    double* out;         // omega_{s+1} (may alias wn in stage 3)
    double* rhs_out;     // optional copy of the RHS
    int Nx, Ny;
    double inv_dx, inv_deta, inv_dx2, inv_deta2, nu, lin_drag;
    const double *a_node, *a_edge, *f_row, *y_node;
    const double* dt_ptr;  // device dt (nullptr: use dt_val)
    double dt_val;
    int mms;
    MmsParams mp;
};

constexpr int SX = 64, SY = 4, RY = 8;

// One thread per column x and RY consecutive rows, with a register window over rows j-1, j, j+1
// of psi and omega_s (columns x-1, x, x+1): each row is loaded once per thread. Interior rows:
// J and the Laplacian from the 3x3 neighbourhoods; wall rows: J = lap = 0 (ops.cu:295, 483).
// The RHS is main.cu's -J + nu lap + f(y) - r omega (+ manufactured forcing), then the SSPRK
// combination (ops.cu:644-655):
//   stage 1: out = ws + dt R;  stage 2: out = 3/4 wn + 1/4 (ws + dt R);
//   stage 3: out = 1/3 wn + 2/3 (ws + dt R)   (out may alias wn: each point reads then writes
//   its own element of wn only)
template <int STAGE, bool MMS>
__global__ void __launch_bounds__(256) stage_kernel(StageArgs a) {
    const int Nx = a.Nx, Ny = a.Ny;
    const int ntx = (Nx + SX - 1) / SX;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);  // row-major tiles, contiguous per XCD
    const int ix = (tile % ntx) * SX + threadIdx.x;
    const int jlo = ((tile / ntx) * SY + threadIdx.y) * RY;
    if (ix >= Nx || jlo >= Ny) return;
    const int jhi = min(jlo + RY, Ny);
    const int xm = wrap_x(ix - 1, Nx), xp = wrap_x(ix + 1, Nx);
    const double inv4dxde = 0.25 * a.inv_dx * a.inv_deta;
    const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
    Stencil9 P, W;  // (m, c, p) of rows j-1 (mm, ym, pm), j (xm, c, xp), j+1 (mp, yp, pp)
    auto row3 = [&](const double* __restrict__ f, int j, double& m, double& c, double& p) {
        const double* r = f + (size_t)j * Nx;
        m = r[xm]; c = r[ix]; p = r[xp];
    };
    if (jlo > 0) {
        row3(a.psi, jlo - 1, P.mm, P.ym, P.pm);
        row3(a.ws, jlo - 1, W.mm, W.ym, W.pm);
    }
    row3(a.psi, jlo, P.xm, P.c, P.xp);
    row3(a.ws, jlo, W.xm, W.c, W.xp);
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        const int j = jlo + q;
        if (j >= jhi) break;
        if (j + 1 < Ny) {
            row3(a.psi, j + 1, P.mp, P.yp, P.pp);
            row3(a.ws, j + 1, W.mp, W.yp, W.pp);
        }
        double J = 0.0, lap = 0.0;
        if (j > 0 && j < Ny - 1) {
            const double an = a.a_node[j];
            double J1, J2, J3;
            arakawa_terms(P, W, an, inv4dxde, J1, J2, J3);
            J = -((J1 + J2 + J3) * (1.0 / 3.0));
            lap = mapped_laplacian(W.xm, W.c, W.xp, W.ym, W.yp, an, a.a_edge[j], a.a_edge[j - 1], a.inv_dx2,
                                   a.inv_deta2);
        }
        const double wc = W.c;
        double rhs = -J + a.nu * lap + a.f_row[j] - a.lin_drag * wc;
        if (MMS) rhs += mms_f_omega(a.mp, ix, a.y_node[j]);
        const size_t c = idx2(j, ix, Nx);
        if (a.rhs_out) a.rhs_out[c] = rhs;
        double o;
        if (STAGE == 1) o = wc + dt * rhs;
        else if (STAGE == 2) o = 0.75 * a.wn[c] + 0.25 * (wc + dt * rhs);
        else o = (1.0 / 3.0) * a.wn[c] + (2.0 / 3.0) * (wc + dt * rhs);
        a.out[c] = o;
        // slide the window down one row
        P.mm = P.xm; P.ym = P.c; P.pm = P.xp; P.xm = P.mp; P.c = P.yp; P.xp = P.pp;
        W.mm = W.xm; W.ym = W.c; W.pm = W.xp; W.xm = W.mp; W.c = W.yp; W.xp = W.pp;
    }
}

// The same stage for Nx a multiple of 64: one wavefront per 64-column strip of RY rows, SYW strips
// stacked in y per workgroup. Each row of psi and omega_s costs one coalesced load per array plus
// a two-lane load of the strip's halo columns (x0-1 for lane 0, x0+64 for lane 63); the x-1 and
// x+1 neighbours come from DPP lane shifts. Rows are fetched PF rows ahead of their use (a ring
// of PF raw rows in registers, and of omega^n rows in stages 2, 3), so that every load has PF
// rows of arithmetic to arrive and each wavefront keeps PF rows of requests in flight.
template <int STAGE, int RY, int SYW, int PF>
__global__ void __launch_bounds__(64 * SYW) stage_kernel_wave(StageArgs a) {
    const int Nx = a.Nx, Ny = a.Ny;
    const int lane = threadIdx.x;
    const int ntx = Nx >> 6;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int x0 = (tile % ntx) << 6, ix = x0 + lane;
    const int jlo = ((tile / ntx) * SYW + threadIdx.y) * RY;
    if (jlo >= Ny) return;  // uniform over the wavefront
    const int jhi = min(jlo + RY, Ny);
    const bool edge = (lane == 0) || (lane == 63);
    const int hx = (lane == 0) ? (x0 == 0 ? Nx - 1 : x0 - 1) : (x0 + 64 == Nx ? 0 : x0 + 64);
    const double inv4dxde = 0.25 * a.inv_dx * a.inv_deta;
    const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
    struct Raw { double pc, ph, wc, wh; };
    auto load = [&](int j, Raw& r) {
        const double* __restrict__ pr = a.psi + (size_t)j * Nx;
        const double* __restrict__ wr = a.ws + (size_t)j * Nx;
        r.pc = pr[ix];
        r.wc = wr[ix];
        r.ph = 0.0;
        r.wh = 0.0;
        if (edge) { r.ph = pr[hx]; r.wh = wr[hx]; }
    };
    // (x-1, x, x+1) of a loaded row
    auto expand = [&](const Raw& r, double& pm, double& pc, double& pp, double& wm, double& wc, double& wp) {
        pm = wave_shr1(r.pc, r.ph); pc = r.pc; pp = wave_shl1(r.pc, r.ph);
        wm = wave_shr1(r.wc, r.wh); wc = r.wc; wp = wave_shl1(r.wc, r.wh);
    };
    Stencil9 P, W;
    Raw ring[PF];     // ring[q % PF] holds row jlo + 1 + q until iteration q uses it
    double wring[PF];  // wring[q % PF] holds omega^n of row jlo + q
#pragma unroll
    for (int f = 0; f < PF; ++f) {
        if (f < RY && jlo + 1 + f < Ny) load(jlo + 1 + f, ring[f]);
        if (STAGE > 1 && f < RY && jlo + f < jhi) wring[f] = a.wn[idx2(jlo + f, ix, Nx)];
    }
    if (jlo > 0) {
        Raw r;
        load(jlo - 1, r);
        expand(r, P.mm, P.ym, P.pm, W.mm, W.ym, W.pm);
    }
    {
        Raw r;
        load(jlo, r);
        expand(r, P.xm, P.c, P.xp, W.xm, W.c, W.xp);
    }
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        const int j = jlo + q;
        if (j >= jhi) break;
        const Raw cur = ring[q % PF];  // row j+1
        const double wn_c = (STAGE > 1) ? wring[q % PF] : 0.0;
        if (q + PF < RY && j + 1 + PF < Ny) load(j + 1 + PF, ring[q % PF]);
        if (STAGE > 1 && q + PF < RY && j + PF < jhi) wring[q % PF] = a.wn[idx2(j + PF, ix, Nx)];
        if (j + 1 < Ny) expand(cur, P.mp, P.yp, P.pp, W.mp, W.yp, W.pp);
        double J = 0.0, lap = 0.0;
        if (j > 0 && j < Ny - 1) {
            const double an = a.a_node[j];
            double J1, J2, J3;
            arakawa_terms(P, W, an, inv4dxde, J1, J2, J3);
            J = -((J1 + J2 + J3) * (1.0 / 3.0));
            lap = mapped_laplacian(W.xm, W.c, W.xp, W.ym, W.yp, an, a.a_edge[j], a.a_edge[j - 1], a.inv_dx2,
                                   a.inv_deta2);
        }
        const double wc = W.c;
        const double rhs = -J + a.nu * lap + a.f_row[j] - a.lin_drag * wc;
        const size_t c = idx2(j, ix, Nx);
        if (a.rhs_out) a.rhs_out[c] = rhs;
        double o;
        if (STAGE == 1) o = wc + dt * rhs;
        else if (STAGE == 2) o = 0.75 * wn_c + 0.25 * (wc + dt * rhs);
        else o = (1.0 / 3.0) * wn_c + (2.0 / 3.0) * (wc + dt * rhs);
        a.out[c] = o;
        P.mm = P.xm; P.ym = P.c; P.pm = P.xp; P.xm = P.mp; P.c = P.yp; P.xp = P.pp;
        W.mm = W.xm; W.ym = W.c; W.pm = W.xp; W.xm = W.mp; W.c = W.yp; W.xp = W.pp;
    }
}

// Forcing profile f(y_j) (ops.cu:526-531), evaluated once per row.
__global__ void forcing_rows_kernel(double* f, const double* y, int Ny, double F0, int nforce, double h) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < Ny) f[j] = kolmogorov_f_omega(F0, nforce, h, y[j]);
}

// Thom closure of both walls plus the enstrophy bookkeeping of main.cu:271-297: the change of
// 1/2 sum_i (w_b omega_b^2 + w_t omega_t^2) caused by the update is added to cum[0].
// One workgroup (the wall rows hold only 2 Nx values).
__global__ void __launch_bounds__(1024) thom_bookkeeping_kernel(double* __restrict__ w,
                                                                const double* __restrict__ psi, int Nx, int Ny,
                                                                const double* __restrict__ dy_edge,
                                                                const double* __restrict__ w_node,
                                                                double psi_bot, double psi_top, int free_slip,
                                                                double* __restrict__ cum) {
    __shared__ double sm[2 * 16];
    double acc[2] = {0.0, 0.0};
    const double qb = w_node[0], qt = w_node[Ny - 1];
    for (int i = threadIdx.x; i < Nx; i += blockDim.x) {
        const double ob = w[i], ot = w[idx2(Ny - 1, i, Nx)];
        double nb, nt;
        thom_values(psi, Nx, Ny, dy_edge, psi_bot, psi_top, free_slip, i, nb, nt);
        w[i] = nb;
        w[idx2(Ny - 1, i, Nx)] = nt;
        acc[0] += qb * ob * ob + qt * ot * ot;
        acc[1] += qb * nb * nb + qt * nt * nt;
    }
    block_sum<2>(acc, sm);
    if (threadIdx.x == 0) *cum += 0.5 * acc[1] - 0.5 * acc[0];
}

// dt rule of main.cu:451-458 from the maxima (umax, vmax, vmaxm) of the current psi.
__global__ void dt_from_maxima_kernel(const double* __restrict__ m, double* __restrict__ dt_out, double dx,
                                      double deta, double cfl, double cvisc, double nu, double inv_dx2,
                                      double inv_dymin2, double dt_max) {
    if (threadIdx.x != 0) return;
    const double umax = m[0], vmaxm = m[2];
    const double dt_x = (umax > 0 ? cfl * (dx / umax) : 1e30);
    const double dt_y = (vmaxm > 0 ? cfl * (deta / vmaxm) : 1e30);
    const double dt_adv = fmin(dt_x, dt_y);
    const double dt_vis = (nu > 0) ? cvisc * 0.5 / (nu * (inv_dx2 + inv_dymin2)) : 1e30;
    *dt_out = fmin(dt_max, fmin(dt_adv, dt_vis));
}

__global__ void advance_time_kernel(double* t, const double* dt) {
    if (threadIdx.x == 0) *t = *t + *dt;
}

}  // namespace

struct ChannelImpl {
    nse2d_channel_desc d;
    hipStream_t s = nullptr;
    double *w = nullptr, *w1 = nullptr, *w2 = nullptr, *psi = nullptr, *rhs = nullptr;
    double *f_row = nullptr, *scalars = nullptr;  // scalars: t, dt, umax, vmax, vmaxm, cum[3]
    PoissonImpl* P = nullptr;
    double inv_dx = 0, inv_deta = 0, inv_dx2 = 0, inv_deta2 = 0, dx = 0, deta = 0;
    // optional timing of kernel classes with events on the launch stream
    int stage_tile = 841;  // wavefront stage kernel shape RY*100 + SYW*10 + PF (0: generic kernel)
    int prof_mask = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> ev_marks;  // (class, index of the start event)
    double prof_ms[NSE2D_PROF_NCLASS] = {0};
    long long prof_n[NSE2D_PROF_NCLASS] = {0};
    double* t_ptr() { return scalars + 0; }
    double* dt_ptr() { return scalars + 1; }
    double* max_ptr() { return scalars + 2; }
    double* cum_ptr() { return scalars + 5; }
};

namespace {

// Brackets a launch of kernel class `cls` with two events when that class is being timed.
struct ProfScope {
    ChannelImpl* c;
    int cls;
    bool on;
    hipEvent_t next_event() {
        if (c->ev_used == c->ev_pool.size()) {
            hipEvent_t e;
            NSE2D_HIP_CHECK(hipEventCreate(&e));
            c->ev_pool.push_back(e);
        }
        return c->ev_pool[c->ev_used++];
    }
    ProfScope(ChannelImpl* ch, int k) : c(ch), cls(k), on((ch->prof_mask >> k) & 1) {
        if (!on) return;
        c->ev_marks.push_back({cls, c->ev_used});
        NSE2D_HIP_CHECK(hipEventRecord(next_event(), c->s));
    }
    ~ProfScope() {
        if (on) NSE2D_HIP_CHECK(hipEventRecord(next_event(), c->s));
    }
};

StageArgs stage_args(ChannelImpl* c, const double* ws, const double* wn, double* out, double* rhs_out,
                     const double* dt_ptr, double dt_val, double t_mms) {
    const nse2d_channel_desc& d = c->d;
    StageArgs a;
    a.psi = c->psi; a.ws = ws; a.wn = wn; a.out = out; a.rhs_out = rhs_out;
    a.Nx = d.Nx; a.Ny = d.Ny;
    a.inv_dx = c->inv_dx; a.inv_deta = c->inv_deta; a.inv_dx2 = c->inv_dx2; a.inv_deta2 = c->inv_deta2;
    a.nu = d.nu; a.lin_drag = d.lin_drag;
    a.a_node = d.d_a_node; a.a_edge = d.d_a_edge; a.f_row = c->f_row; a.y_node = d.d_y_node;
    a.dt_ptr = dt_ptr; a.dt_val = dt_val;
    a.mms = d.mms;
    if (d.mms) a.mp = make_mms_params(d.Nx, d.Lx, d.h, t_mms, d.nu, d.lin_drag, d.mms_A0, d.mms_eps, d.mms_Om, d.mms_kx);
    else a.mp = MmsParams{};
    return a;
}

// Shape of the wavefront stage kernel: RY rows per wavefront, SYW wavefronts per workgroup and
// prefetch depth PF, coded RY*100 + SYW*10 + PF. NSE2D_STAGE_TILE=RYxSYWxPF (or RYxSYW, PF = 2)
// selects another instantiated shape (used to tune on the device); 0 selects the generic kernel.
// The manufactured-solution forcing (sin/cos per point) always runs in the generic kernel, so the
// production kernels carry none of its registers.
constexpr int kStageDefault = 841;
int stage_tile_code() {
    int ry = 8, sy = 4, pf = 1;
    if (const char* e = std::getenv("NSE2D_STAGE_TILE")) {
        if (e[0] == '0') return 0;
        const int n = std::sscanf(e, "%dx%dx%d", &ry, &sy, &pf);
        if (n == 2) pf = 2;
        else if (n != 3) return kStageDefault;
    }
    const int c = ry * 100 + sy * 10 + pf;
    switch (c) {
        case 841: case 842: case 843: case 1641: case 1642: case 1643: case 1621: case 1622: case 1623:
        case 3211: case 3212: case 3213: return c;
        default: return kStageDefault;
    }
}

template <int STAGE, bool MMS>
void launch_stage_t(ChannelImpl* c, const StageArgs& a) {
    const int code = (!MMS && c->d.Nx % 64 == 0) ? c->stage_tile : 0;
#define NSE2D_STAGE_WAVE(RYV, SYV, PFV)                                                                  \
    case RYV * 100 + SYV * 10 + PFV:                                                                     \
        stage_kernel_wave<STAGE, RYV, SYV, PFV><<<(c->d.Nx / 64) * ceil_div(c->d.Ny, RYV * SYV),         \
                                                  dim3(64, SYV), 0, c->s>>>(a);                          \
        break;
    switch (code) {
        NSE2D_STAGE_WAVE(8, 4, 1) NSE2D_STAGE_WAVE(8, 4, 2) NSE2D_STAGE_WAVE(8, 4, 3)
        NSE2D_STAGE_WAVE(16, 4, 1) NSE2D_STAGE_WAVE(16, 4, 2) NSE2D_STAGE_WAVE(16, 4, 3)
        NSE2D_STAGE_WAVE(16, 2, 1) NSE2D_STAGE_WAVE(16, 2, 2) NSE2D_STAGE_WAVE(16, 2, 3)
        NSE2D_STAGE_WAVE(32, 1, 1) NSE2D_STAGE_WAVE(32, 1, 2) NSE2D_STAGE_WAVE(32, 1, 3)
        default: {
            const unsigned ntiles = ceil_div(c->d.Nx, SX) * ceil_div(c->d.Ny, SY * RY);
            stage_kernel<STAGE, MMS><<<ntiles, dim3(SX, SY), 0, c->s>>>(a);
        }
    }
#undef NSE2D_STAGE_WAVE
}

template <int STAGE>
void launch_stage(ChannelImpl* c, const StageArgs& a) {
    ProfScope ps(c, NSE2D_PROF_STAGE);
    if (a.mms) launch_stage_t<STAGE, true>(c, a);
    else launch_stage_t<STAGE, false>(c, a);
    NSE2D_LAUNCH_CHECK();
}

void close_stage(ChannelImpl* c, double* f, int stage, bool bookkeeping) {
    const nse2d_channel_desc& d = c->d;
    { ProfScope ps(c, NSE2D_PROF_FFT_FWD); poisson_forward(c->P, f, c->s); }
    { ProfScope ps(c, NSE2D_PROF_THOMAS); poisson_tridiag(c->P, d.psi_bot, d.psi_top, c->s); }
    if (!bookkeeping && poisson_fuses_closure(c->P)) {
        // the closure rides in the inverse transform's workgroups of rows 1 and Ny-2
        ProfScope ps(c, NSE2D_PROF_FFT_INV);
        poisson_inverse_closure(c->P, c->psi, f, d.d_dy_edge, d.psi_bot, d.psi_top, wall_type(), c->s);
        return;
    }
    { ProfScope ps(c, NSE2D_PROF_FFT_INV); poisson_inverse(c->P, c->psi, c->s); }
    ProfScope ps(c, NSE2D_PROF_THOM);
    if (bookkeeping) {
        thom_bookkeeping_kernel<<<1, 1024, 0, c->s>>>(f, c->psi, d.Nx, d.Ny, d.d_dy_edge, d.d_w_node, d.psi_bot,
                                                      d.psi_top, wall_type(), c->cum_ptr() + stage);
        NSE2D_LAUNCH_CHECK();
    } else {
        launch_thom(f, c->psi, d.Nx, d.Ny, d.d_dy_edge, d.psi_bot, d.psi_top, wall_type(), c->s);
    }
}

void do_step(ChannelImpl* c, double t, double dt, const double* dt_dev, int flags) {
    const bool book = (flags & NSE2D_STEP_WALL_BOOKKEEPING) != 0;
    double* rhs3 = (flags & NSE2D_STEP_STORE_RHS) ? c->rhs : nullptr;
    launch_stage<1>(c, stage_args(c, c->w, nullptr, c->w1, nullptr, dt_dev, dt, t));
    close_stage(c, c->w1, 0, book);
    launch_stage<2>(c, stage_args(c, c->w1, c->w, c->w2, nullptr, dt_dev, dt, t + dt));
    close_stage(c, c->w2, 1, book);
    launch_stage<3>(c, stage_args(c, c->w2, c->w, c->w, rhs3, dt_dev, dt, t + 0.5 * dt));
    close_stage(c, c->w, 2, book);
}

}  // namespace

}  // namespace nse2d

struct nse2d_channel {
    nse2d::ChannelImpl impl;
};

using namespace nse2d;

extern "C" {

nse2d_channel_t nse2d_channel_create(const nse2d_channel_desc* desc, hipStream_t s) {
    if (!desc) { set_last_error("nse2d_channel_create: null descriptor"); return nullptr; }
    const nse2d_channel_desc& d = *desc;
    if (!d.d_a_node || !d.d_a_edge || !d.d_y_node || !d.d_dy_edge || !d.d_w_node) {
        set_last_error("nse2d_channel_create: metric arrays must be set");
        return nullptr;
    }
    PoissonImpl* P = poisson_create(d.Nx, d.Ny, d.Lx, s);
    if (!P) return nullptr;
    auto* ch = new nse2d_channel;
    ChannelImpl* c = &ch->impl;
    c->d = d;
    c->s = s;
    c->P = P;
    c->stage_tile = stage_tile_code();
    c->dx = d.Lx / (double)d.Nx;
    c->inv_dx = 1.0 / c->dx;
    c->inv_dx2 = c->inv_dx * c->inv_dx;
    c->deta = 2.0 / (double)(d.Ny - 1);
    c->inv_deta = 1.0 / c->deta;
    c->inv_deta2 = c->inv_deta * c->inv_deta;
    poisson_enable_mapped(P, d.d_a_node, d.d_a_edge, c->inv_deta, s);
    const size_t N = (size_t)d.Nx * d.Ny;
    for (double** f : {&c->w, &c->w1, &c->w2, &c->psi, &c->rhs}) {
        NSE2D_HIP_CHECK(hipMalloc(f, N * sizeof(double)));
        NSE2D_HIP_CHECK(hipMemsetAsync(*f, 0, N * sizeof(double), s));
    }
    NSE2D_HIP_CHECK(hipMalloc(&c->f_row, d.Ny * sizeof(double)));
    NSE2D_HIP_CHECK(hipMalloc(&c->scalars, 16 * sizeof(double)));
    NSE2D_HIP_CHECK(hipMemsetAsync(c->scalars, 0, 16 * sizeof(double), s));
    forcing_rows_kernel<<<ceil_div(d.Ny, 256), 256, 0, s>>>(c->f_row, d.d_y_node, d.Ny, d.F0, d.nforce, d.h);
    NSE2D_LAUNCH_CHECK();
    NSE2D_HIP_CHECK(hipStreamSynchronize(s));
    return ch;
}

void nse2d_channel_destroy(nse2d_channel_t ch) {
    if (!ch) return;
    ChannelImpl* c = &ch->impl;
    (void)hipStreamSynchronize(c->s);
    for (double* f : {c->w, c->w1, c->w2, c->psi, c->rhs, c->f_row, c->scalars}) (void)hipFree(f);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    poisson_destroy(c->P);
    delete ch;
}

double* nse2d_channel_field(nse2d_channel_t ch, int which) {
    ChannelImpl* c = &ch->impl;
    switch (which) {
        case NSE2D_FIELD_OMEGA: return c->w;
        case NSE2D_FIELD_PSI: return c->psi;
        case NSE2D_FIELD_RHS: return c->rhs;
        case NSE2D_FIELD_W1: return c->w1;
        case NSE2D_FIELD_W2: return c->w2;
        default: return nullptr;
    }
}

void nse2d_channel_init(nse2d_channel_t ch) {
    ChannelImpl* c = &ch->impl;
    const nse2d_channel_desc& d = c->d;
    poisson_solve(c->P, c->w, c->psi, d.psi_bot, d.psi_top, c->s);
    launch_thom(c->w, c->psi, d.Nx, d.Ny, d.d_dy_edge, d.psi_bot, d.psi_top, wall_type(), c->s);
}

void nse2d_channel_step(nse2d_channel_t ch, double t, double dt, int flags) {
    do_step(&ch->impl, t, dt, nullptr, flags);
}

void nse2d_channel_step_auto(nse2d_channel_t ch, int n) {
    ChannelImpl* c = &ch->impl;
    const nse2d_channel_desc& d = c->d;
    if (d.mms) {
        std::fprintf(stderr, "nse2d_channel_step_auto: manufactured forcing needs host time\n");
        std::abort();
    }
    const double inv_dymin2 = 1.0 / (d.dy_min * d.dy_min);
    for (int k = 0; k < n; ++k) {
        double* m = c->max_ptr();
        {
            ProfScope ps(c, NSE2D_PROF_CFL);
            launch_max_uv(c->psi, d.Nx, d.Ny, c->inv_dx, c->inv_deta, d.d_a_node, m, m + 1, m + 2, c->s);
            dt_from_maxima_kernel<<<1, 64, 0, c->s>>>(m, c->dt_ptr(), c->dx, c->deta, d.cfl, d.cvisc, d.nu,
                                                      c->inv_dx2, inv_dymin2, d.dt_max);
            NSE2D_LAUNCH_CHECK();
        }
        do_step(c, 0.0, 0.0, c->dt_ptr(), 0);
        advance_time_kernel<<<1, 64, 0, c->s>>>(c->t_ptr(), c->dt_ptr());
        NSE2D_LAUNCH_CHECK();
    }
}

void nse2d_channel_profile(nse2d_channel_t ch, int class_mask) {
    ChannelImpl* c = &ch->impl;
    c->prof_mask = class_mask;
    for (int k = 0; k < NSE2D_PROF_NCLASS; ++k) { c->prof_ms[k] = 0; c->prof_n[k] = 0; }
    c->ev_used = 0;
    c->ev_marks.clear();
}

void nse2d_channel_profile_read(nse2d_channel_t ch, double* ms, long long* count) {
    ChannelImpl* c = &ch->impl;
    NSE2D_HIP_CHECK(hipStreamSynchronize(c->s));
    for (auto& m : c->ev_marks) {
        float dt_ms = 0.f;
        NSE2D_HIP_CHECK(hipEventElapsedTime(&dt_ms, c->ev_pool[m.second], c->ev_pool[m.second + 1]));
        c->prof_ms[m.first] += dt_ms;
        c->prof_n[m.first] += 1;
    }
    c->ev_used = 0;
    c->ev_marks.clear();
    for (int k = 0; k < NSE2D_PROF_NCLASS; ++k) {
        if (ms) ms[k] = c->prof_ms[k];
        if (count) count[k] = c->prof_n[k];
    }
}

double* nse2d_channel_time_ptr(nse2d_channel_t ch) { return ch->impl.t_ptr(); }
double* nse2d_channel_dt_ptr(nse2d_channel_t ch) { return ch->impl.dt_ptr(); }

void nse2d_channel_wall_bookkeeping(nse2d_channel_t ch, double out3[3], double invA_full) {
    ChannelImpl* c = &ch->impl;
    double raw[3];
    NSE2D_HIP_CHECK(hipMemcpyAsync(raw, c->cum_ptr(), 3 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(c->s));
    for (int k = 0; k < 3; ++k) out3[k] = raw[k] * invA_full;
}

}  // extern "C"
