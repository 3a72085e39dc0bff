// The SSPRK(3,3) step of the reference driver (main.cu:482-538) on the device:
//   for s = 1..3:  stage kernel (Arakawa J + mapped Laplacian + RHS + forcing + update, fused)
//                  -> Poisson solve -> Thom wall closure
// plus the device-side adaptive dt (main.cu:440-458) so that steps need no host round trip.
#include "nse2d_common.hpp"
#include "nse2d_ops.hpp"
#include "nse2d_poisson.hpp"
#include "fft_rows.hpp"
#include "../../include/nse2d_hip.h"

#include <algorithm>
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
    const float* closure;  // optional additive RHS term of the interior rows ([Ny][Nx]), or nullptr
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
        if (a.closure && j > 0 && j < Ny - 1) rhs += (double)a.closure[c];
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
__device__ __forceinline__ void stage_wave_body(const StageArgs& a, int tile) {
    const int Nx = a.Nx, Ny = a.Ny;
    const int lane = threadIdx.x;
    const int ntx = Nx >> 6;
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

// The same stage for Nx a multiple of 128 with two columns per lane: a wavefront owns a 128-column
// strip, lane l holds columns x0 + 2l and x0 + 2l + 1, and every row moves as one 16-byte access
// per lane (1 KB per wavefront instruction). The x-neighbours outside the pair come from the DPP
// shifts of the other element (lane l-1's right value, lane l+1's left value), the strip's halo
// columns from lanes 0 and 63.
// One row of a 128-column strip as loaded: the lane's pair of psi and omega_s, and (lanes 0 and
// 63) the strip's halo column.
struct PairRow { double2 p, w; double ph, wh; };
// (x0+2l-1, x0+2l, x0+2l+1, x0+2l+2) of a loaded row
struct Row4 { double v[4]; };

// Row j's pair of columns (ix, ix+1) of psi and omega_s, and the column hx beside the strip (used
// by the edge lanes; the others all load one line). No load is conditional: under a branch the
// compiler waits, at the join, for as many loads as the path without them needs, and since loads
// complete in order that included the rows prefetched after them.
__device__ __forceinline__ void load_pair_row(const StageArgs& a, int j, int ix, int hx, PairRow& r) {
    const double* __restrict__ pr = a.psi + (size_t)j * a.Nx;
    const double* __restrict__ wr = a.ws + (size_t)j * a.Nx;
    r.p = *reinterpret_cast<const double2*>(pr + ix);
    r.w = *reinterpret_cast<const double2*>(wr + ix);
    r.ph = pr[hx];
    r.wh = wr[hx];
}
__device__ __forceinline__ void expand_pair_row(const PairRow& r, Row4& P4, Row4& W4) {
    P4.v[0] = wave_shr1(r.p.y, r.ph); P4.v[1] = r.p.x; P4.v[2] = r.p.y; P4.v[3] = wave_shl1(r.p.x, r.ph);
    W4.v[0] = wave_shr1(r.w.y, r.wh); W4.v[1] = r.w.x; W4.v[2] = r.w.y; W4.v[3] = wave_shl1(r.w.x, r.wh);
}
__device__ __forceinline__ Stencil9 stencil_at(const Row4& m, const Row4& c, const Row4& pl, int e) {
    Stencil9 S;
    S.xm = c.v[e]; S.c = c.v[e + 1]; S.xp = c.v[e + 2];
    S.mm = m.v[e]; S.ym = m.v[e + 1]; S.pm = m.v[e + 2];
    S.mp = pl.v[e]; S.yp = pl.v[e + 1]; S.pp = pl.v[e + 2];
    return S;
}

// Row j's metric factors and forcing (zero metrics on the wall rows, which take no J or Laplacian).
struct RowCoef { double an, ae, aw, fr; };
__device__ __forceinline__ RowCoef row_coef(const StageArgs& a, int j) {
    const bool interior = (j > 0 && j < a.Ny - 1);
    return {interior ? a.a_node[j] : 0.0, interior ? a.a_edge[j] : 0.0, interior ? a.a_edge[j - 1] : 0.0,
            a.f_row[j]};
}
// The same with unconditional loads (indices clamped into the arrays, wall rows then select
// zero), for the sweeps that load each row's factors with the rows' vector loads: a load under a
// branch makes the compiler wait, at the join, for the loads issued after it too.
__device__ __forceinline__ RowCoef row_coef_loaded(const StageArgs& a, int j) {
    const bool interior = (j > 0 && j < a.Ny - 1);
    const double an = a.a_node[j], ae = a.a_edge[min(j, a.Ny - 2)], aw = a.a_edge[max(j - 1, 0)];
    return {interior ? an : 0.0, interior ? ae : 0.0, interior ? aw : 0.0, a.f_row[j]};
}

// The two updated values (columns ix, ix+1) of row j from the expanded rows j-1, j, j+1 of psi
// and omega_s and the lane's pair of omega^n (stages 2, 3); stores the RHS copy to rhs_out when
// it is not null.
template <int STAGE>
__device__ __forceinline__ double2 stage_pair(const StageArgs& a, int j, int ix, const Row4& Pm, const Row4& Pc,
                                              const Row4& Pp, const Row4& Wm, const Row4& Wc, const Row4& Wp,
                                              double2 wn_c, double dt, double inv4dxde, const RowCoef& rk,
                                              double* __restrict__ rhs_out) {
    double o[2];
    const bool interior = (j > 0 && j < a.Ny - 1);
    const double an = rk.an, ae = rk.ae, aw = rk.aw, fr = rk.fr;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        double J = 0.0, lap = 0.0;
        const Stencil9 P = stencil_at(Pm, Pc, Pp, e), W = stencil_at(Wm, Wc, Wp, e);
        if (interior) {
            double J1, J2, J3;
            arakawa_terms(P, W, an, inv4dxde, J1, J2, J3);
            J = -((J1 + J2 + J3) * (1.0 / 3.0));
            lap = mapped_laplacian(W.xm, W.c, W.xp, W.ym, W.yp, an, ae, aw, a.inv_dx2, a.inv_deta2);
        }
        const double wc = W.c;
        const double rhs = -J + a.nu * lap + fr - a.lin_drag * wc;
        if (rhs_out) rhs_out[idx2(j, ix + e, a.Nx)] = rhs;
        const double wn = e ? wn_c.y : wn_c.x;
        if (STAGE == 1) o[e] = wc + dt * rhs;
        else if (STAGE == 2) o[e] = 0.75 * wn + 0.25 * (wc + dt * rhs);
        else o[e] = (1.0 / 3.0) * wn + (2.0 / 3.0) * (wc + dt * rhs);
    }
    return double2{o[0], o[1]};
}

// One wavefront's RY rows [jlo, jhi) of a 128-column strip, swept downwards (increasing j) or,
// UP, upwards. ROWS_FIRST: the row factors of all RY rows are loaded at the start (scalar loads
// into SGPRs when the arguments are kernel arguments; the group kernels, whose arguments come
// from device memory, would hold them in VGPRs and keep the row-by-row loads).
// rk: the row factors of rows jlo + q (ROWS_FIRST); an upward sweep covers a whole slab (jhi =
// jlo + RY), so that it reads them at fixed indices.
template <int STAGE, int RY, int PF, bool ROWS_FIRST, bool UP>
__device__ __forceinline__ void stage_wave2_sweep(const StageArgs& a, int ix, int jlo, int jhi, int hx,
                                                  const RowCoef (&rk)[ROWS_FIRST ? RY : 1]) {
    const int Nx = a.Nx, Ny = a.Ny;
    const double inv4dxde = 0.25 * a.inv_dx * a.inv_deta;
    const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
    // step q updates row rowq(q); its new neighbour is row rowq(q) + dj
    constexpr int dj = UP ? -1 : 1;
    auto rowq = [&](int q) { return UP ? jhi - 1 - q : jlo + q; };
    // Every load is unconditional: rows beyond the walls are clamped to the wall row and rows past
    // a short last slab to its last row (the wall rows take no stencil, and nothing reads the
    // others). Under a branch, the compiler waited at the join for as many loads as the path
    // without them needs, which (loads complete in order) included the row prefetched ahead.
    auto wall = [&](int j) { return min(max(j, 0), Ny - 1); };
    // without ROWS_FIRST, each row's factors are loaded a row ahead, with the row's prefetch
    RowCoef rkn = ROWS_FIRST ? RowCoef{} : row_coef_loaded(a, rowq(0));
    auto row_factors = [&](int q, int j) {
        if constexpr (ROWS_FIRST) return rk[UP ? RY - 1 - q : q];
        else return rkn;
    };
    const int n = jhi - jlo;
    // Pb: the row behind the sweep (j - dj), Pc: row j, Pf: the row ahead (j + dj)
    Row4 Pb{}, Pc{}, Pf{}, Wb{}, Wc{}, Wf{};
    PairRow ring[PF];
    double2 wring[PF];
#pragma unroll
    for (int f = 0; f < PF; ++f) {
        if (f < RY) load_pair_row(a, wall(rowq(0) + dj * (1 + f)), ix, hx, ring[f]);
        if (STAGE > 1 && f < RY) wring[f] = *reinterpret_cast<const double2*>(a.wn + idx2(rowq(min(f, n - 1)), ix, Nx));
    }
    {
        PairRow r;
        load_pair_row(a, wall(rowq(0) - dj), ix, hx, r);
        expand_pair_row(r, Pb, Wb);
    }
    {
        PairRow r;
        load_pair_row(a, rowq(0), ix, hx, r);
        expand_pair_row(r, Pc, Wc);
    }
#pragma unroll
    for (int q = 0; q < RY; ++q) {
        if (q >= n) break;
        const int j = rowq(q);
        const PairRow cur = ring[q % PF];  // row j + dj
        const double2 wn_c = (STAGE > 1) ? wring[q % PF] : double2{0.0, 0.0};
        const RowCoef rq = row_factors(q, j);
        if (q + PF < RY) load_pair_row(a, wall(j + dj * (1 + PF)), ix, hx, ring[q % PF]);
        if (STAGE > 1 && q + PF < RY)
            wring[q % PF] = *reinterpret_cast<const double2*>(a.wn + idx2(rowq(min(q + PF, n - 1)), ix, Nx));
        if (!ROWS_FIRST && q + 1 < RY) rkn = row_coef_loaded(a, rowq(min(q + 1, n - 1)));
        expand_pair_row(cur, Pf, Wf);
        const double2 o = UP ? stage_pair<STAGE>(a, j, ix, Pf, Pc, Pb, Wf, Wc, Wb, wn_c, dt, inv4dxde, rq,
                                                 a.rhs_out)
                             : stage_pair<STAGE>(a, j, ix, Pb, Pc, Pf, Wb, Wc, Wf, wn_c, dt, inv4dxde, rq,
                                                 a.rhs_out);
        *reinterpret_cast<double2*>(a.out + idx2(j, ix, Nx)) = o;
        Pb = Pc; Pc = Pf; Wb = Wc; Wc = Wf;
    }
}

// The wavefronts' RY-row slabs alternate in direction: even slabs are swept downwards, odd ones
// upwards, so the two rows that neighbouring slabs both read are read by both at the same time
// (both at their first rows or both at their last), while the lines are in L2. Swept in one
// direction, a slab reads the row it shares with the slab above first and that slab reads it
// last, a whole slab's time later: at 4096x2049 one read in five of psi and omega_s then came
// from HBM again (3.5 instead of 3 fields per stage).
template <int STAGE, int RY, int SYW, int PF, bool ROWS_FIRST = true>
__device__ __forceinline__ void stage_wave2_body(const StageArgs& a, int tile) {
    const int Nx = a.Nx, Ny = a.Ny;
    const int lane = threadIdx.x;
    const int ntx = Nx >> 7;
    const int x0 = (tile % ntx) << 7, ix = x0 + 2 * lane;
    // the wavefront's slab and first row (uniform over the wavefront: its row factors are scalar loads)
    const int slab = __builtin_amdgcn_readfirstlane((tile / ntx) * SYW + (int)threadIdx.y);
    const int jlo = slab * RY;
    if (jlo >= Ny) return;  // uniform over the wavefront
    const int jhi = min(jlo + RY, Ny);
    const int hx = (lane == 0) ? (x0 == 0 ? Nx - 1 : x0 - 1) : (x0 + 128 == Nx ? 0 : x0 + 128);
    // The row factors of all RY rows, loaded first. Loaded row by row, after the previous row's
    // store (which the compiler cannot tell apart from them), each would wait behind the row
    // prefetched ahead of it, and every row would pay a full memory latency.
    RowCoef rk[ROWS_FIRST ? RY : 1];
    if constexpr (ROWS_FIRST) {
#pragma unroll
        for (int q = 0; q < RY; ++q) rk[q] = row_coef(a, min(jlo + q, Ny - 1));
    }
    // a partial last slab is swept downwards
    if ((slab & 1) && jhi == jlo + RY)
        stage_wave2_sweep<STAGE, RY, PF, ROWS_FIRST, true>(a, ix, jlo, jhi, hx, rk);
    else
        stage_wave2_sweep<STAGE, RY, PF, ROWS_FIRST, false>(a, ix, jlo, jhi, hx, rk);
}

// The stage and the forward row transform of the Poisson solve in one kernel, one workgroup per
// grid row (rows contiguous per XCD, so that the rows j-1 and j+1 each workgroup reads are the
// neighbouring workgroups' rows and meet in that XCD's L2). Thread tid holds the transform's
// first-pass inputs z_{tid + q T} = (x_{2(tid + q T)}, x_{2(tid + q T) + 1}), q < R: in slot q the
// wavefronts cover 128-column strips, so each slot is one row of the two-column stage (same
// arithmetic, stage_pair). The updated row is stored, and on interior rows it goes from
// registers through the row transform into the spectrum: the solve's forward phase then needs
// no second read of omega_{s+1}. Production stages only (no manufactured forcing, closure term
// or RHS copy).
template <int STAGE, int M, int R>
__global__ void __launch_bounds__(M / R) stage_fft_kernel(StageArgs a, fft::cplx* __restrict__ hat,
                                                          const fft::cplx* __restrict__ twp,
                                                          const fft::cplx* __restrict__ tw) {
    constexpr int T = M / R;
    __shared__ fft::cplx lds[fft::Plan<M, R>::LDS];
    const int Nx = 2 * M, Ny = a.Ny;
    const int j = xcd_tile(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x, lane = tid & 63;
    const bool interior = j > 0 && j < Ny - 1;  // uniform over the workgroup
    const double inv4dxde = 0.25 * a.inv_dx * a.inv_deta;
    const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
    // the row factors, loaded once (the compiler cannot tell that the stores below leave them alone)
    const RowCoef rk = row_coef(a, j);
    // slot q: the wavefront's 128-column strip x0 = 2 (q T + 64 w)
    auto strip = [&](int q) { return 2 * (q * T + (tid & ~63)); };
    auto halo = [&](int x0) { return (lane == 0) ? (x0 == 0 ? Nx - 1 : x0 - 1) : (x0 + 128 == Nx ? 0 : x0 + 128); };
    struct Slot { PairRow m, c, p; double2 wn; };
    auto load_slot = [&](int q, Slot& S) {
        const int x0 = strip(q), ix = x0 + 2 * lane, hx = halo(x0);
        load_pair_row(a, j, ix, hx, S.c);
        if (interior) {
            load_pair_row(a, j - 1, ix, hx, S.m);
            load_pair_row(a, j + 1, ix, hx, S.p);
        }
        S.wn = (STAGE > 1) ? *reinterpret_cast<const double2*>(a.wn + idx2(j, ix, Nx)) : double2{0.0, 0.0};
    };
    fft::cplx v[R];
    Slot ring[2];
    load_slot(0, ring[0]);
#pragma unroll
    for (int q = 0; q < R; ++q) {
        // the next slot's loads are in flight while this slot is computed
        if (q + 1 < R) load_slot(q + 1, ring[(q + 1) & 1]);
        const Slot& S = ring[q & 1];
        const int ix = strip(q) + 2 * lane;
        Row4 Pm{}, Pc{}, Pp{}, Wm{}, Wc{}, Wp{};
        expand_pair_row(S.c, Pc, Wc);
        if (interior) {
            expand_pair_row(S.m, Pm, Wm);
            expand_pair_row(S.p, Pp, Wp);
        }
        const double2 o = stage_pair<STAGE>(a, j, ix, Pm, Pc, Pp, Wm, Wc, Wp, S.wn, dt, inv4dxde, rk, nullptr);
        *reinterpret_cast<double2*>(a.out + idx2(j, ix, Nx)) = o;
        v[q] = {o.x, o.y};
    }
    if (interior) fft::forward_row_from<M, R>(v, hat + (size_t)j * M, lds, twp, tw);
}

template <int STAGE, int RY, int SYW, int PF>
__global__ void __launch_bounds__(64 * SYW) stage_kernel_wave2(StageArgs a) {
    // (the row factors of more than 8 rows do not fit the scalar registers: loaded row by row)
    stage_wave2_body<STAGE, RY, SYW, PF, (RY <= 8)>(a, xcd_tile(blockIdx.x, gridDim.x));
}

template <int STAGE, int RY, int SYW, int PF>
__global__ void __launch_bounds__(64 * SYW) stage_kernel_wave(StageArgs a) {
    stage_wave_body<STAGE, RY, SYW, PF>(a, xcd_tile(blockIdx.x, gridDim.x));  // tiles contiguous per XCD
}

// The member whose range [begin(m), begin(m + 1)) holds the global tile b (begin ascending,
// begin(0) = 0): a binary search, log2(n) dependent loads before the tile's own loads (a linear
// walk over 64 members cost up to 63).
template <class Begin>
__device__ __forceinline__ int member_of(int b, int n, Begin begin) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin(mid) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The stage of several channels (ensemble members) in one launch: args[m] describes member m and
// its tiles are the global tiles [tile_begin[m], tile_begin[m+1]); the XCD renumbering runs over
// the global tile index, so each XCD works through one contiguous range of members' rows.
template <int STAGE, int RY, int SYW, int PF>
__global__ void __launch_bounds__(64 * SYW) stage_kernel_wave_group(const StageArgs* __restrict__ args,
                                                                    const int* __restrict__ tile_begin, int n) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int m = member_of(b, n, [&](int k) { return tile_begin[k]; });
    stage_wave_body<STAGE, RY, SYW, PF>(args[m], b - tile_begin[m]);
}

// The two-column stage of several channels in one launch (tiles as stage_kernel_wave_group, of
// 128-column strips).
template <int STAGE, int RY, int SYW, int PF>
__global__ void __launch_bounds__(64 * SYW) stage_kernel_wave2_group(const StageArgs* __restrict__ args,
                                                                     const int* __restrict__ tile_begin, int n) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int m = member_of(b, n, [&](int k) { return tile_begin[k]; });
    stage_wave2_body<STAGE, RY, SYW, PF, false>(args[m], b - tile_begin[m]);
}

// Persistent form of the grouped two-column stage: a fixed grid (a few workgroups per CU) walks
// all members' tiles. Each XCD owns the same contiguous range of global tiles as under xcd_tile,
// and its workgroups take that range in order, gx tiles apart, so that neighbouring tiles are in
// flight on one XCD at the same time.
template <int STAGE, int RY, int SYW, int PF>
__global__ void __launch_bounds__(64 * SYW) stage_kernel_wave2_group_persistent(const StageArgs* __restrict__ args,
                                                                                const int* __restrict__ tile_begin,
                                                                                int n, int total) {
    const int x = blockIdx.x % kXcds, l = blockIdx.x / kXcds, G = gridDim.x;
    const int gx = G / kXcds + (x < G % kXcds ? 1 : 0);  // workgroups on this XCD
    const int q = total / kXcds, rr = total % kXcds;
    const int t0 = x * q + min(x, rr), t1 = t0 + q + (x < rr ? 1 : 0);
    int m = 0;
    for (int b = t0 + l; b < t1; b += gx) {
        while (m + 1 < n && tile_begin[m + 1] <= b) ++m;  // b only grows: the member index only advances
        stage_wave2_body<STAGE, RY, SYW, PF, false>(args[m], b - tile_begin[m]);
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
struct DtRule {
    double dx, deta, cfl, cvisc, nu, inv_dx2, inv_dymin2, dt_max;
};
__device__ __forceinline__ double dt_rule(const double* m, const DtRule& r) {
    const double umax = m[0], vmaxm = m[2];
    const double dt_x = (umax > 0 ? r.cfl * (r.dx / umax) : 1e30);
    const double dt_y = (vmaxm > 0 ? r.cfl * (r.deta / vmaxm) : 1e30);
    const double dt_adv = fmin(dt_x, dt_y);
    const double dt_vis = (r.nu > 0) ? r.cvisc * 0.5 / (r.nu * (r.inv_dx2 + r.inv_dymin2)) : 1e30;
    return fmin(r.dt_max, fmin(dt_adv, dt_vis));
}
__global__ void dt_from_maxima_kernel(const double* __restrict__ m, double* __restrict__ dt_out, DtRule r) {
    if (threadIdx.x == 0) *dt_out = dt_rule(m, r);
}

// The CFL maxima of psi (diag.cu:702-866), the dt rule (main.cu:451-458) and the time advance in
// two launches: Ny workgroups reduce one row each to its three maxima (cfl_rows_kernel), then one
// workgroup reduces the Ny partials in row order and writes the maxima (scal[2..4]), dt (scal[1])
// and t + dt (scal[0]) (cfl_final_kernel). A single launch that lets its last workgroup finish
// needs a device-scope release fence in every workgroup, which writes back the XCD's L2 each time
// (measured 2x slower than the two launches).
constexpr int kCflThreads = 256;
__global__ void __launch_bounds__(kCflThreads) cfl_rows_kernel(const double* __restrict__ psi, int Nx, int Ny,
                                                               double inv_dx, double inv_deta,
                                                               const double* __restrict__ a_node,
                                                               double* __restrict__ partial) {
    __shared__ double sm[3 * 16];
    row_max_uv_body(psi, Nx, Ny, inv_dx, inv_deta, a_node, xcd_tile(blockIdx.x, gridDim.x), partial, sm);
}

// The same maxima for Nx a multiple of 512: workgroup tiles of 512 columns x kCflRows interior rows
// (rows 1 + kCflRows band ..), one 128-column strip per wavefront, lane l holding columns x0 + 2l
// and x0 + 2l + 1 (16-byte loads); each wavefront walks down its rows with a window of three rows
// in registers and the row after next already in flight. x-neighbours come from DPP shifts and the
// strip's two halo columns. Workgroup b writes its maxima to partial[a * gridDim.x + b].
constexpr int kCflRows = 16;
// Tile b of ntiles of one field; partial[a * ntiles + b] receives maximum a.
__device__ __forceinline__ void cfl_tile_body(const double* __restrict__ psi, int Nx, int Ny, double inv_dx,
                                              double inv_deta, const double* __restrict__ a_node, int b, int ntiles,
                                              double* __restrict__ partial, double* sm) {
    const int nsx = Nx / 512;
    const int sx = b % nsx, band = b / nsx;
    const int j0 = 1 + band * kCflRows, j1 = min(j0 + kCflRows, Ny - 1);  // rows [j0, j1)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x0 = sx * 512 + wave * 128, ix = x0 + 2 * lane;
    const int hx = (lane == 0) ? (x0 == 0 ? Nx - 1 : x0 - 1) : (x0 + 128 == Nx ? 0 : x0 + 128);
    const double hdy = 0.5 * inv_deta, hdx = 0.5 * inv_dx;
    auto row = [&](int j) { return *reinterpret_cast<const double2*>(psi + idx2(j, ix, Nx)); };
    // the column beside the strip (used by the edge lanes only; every lane loads it, the others
    // all from one line)
    auto halo = [&](int j) { return psi[idx2(j, hx, Nx)]; };
    // rows j0 - 1 .. j1 in rr (row j0 - 1 + i in rr[i]), halos of rows j0 .. j1 - 1 in hh; each
    // is issued PD rows before the row that first reads it. The loads are unconditional (clamped
    // at the band's end, values not used) and the loop is unrolled: with the loads under a branch,
    // or in a rolled loop, the compiler waited for each row as soon as it was issued (loads
    // complete in order).
    constexpr int PD = 4;
    double2 rr[kCflRows + 2];
    double hh[kCflRows];
#pragma unroll
    for (int i = 0; i < PD + 2; ++i) rr[i] = row(min(j0 - 1 + i, j1));
#pragma unroll
    for (int i = 0; i < PD; ++i) hh[i] = halo(min(j0 + i, j1 - 1));
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
#pragma unroll
    for (int q = 0; q < kCflRows; ++q) {
        const int j = j0 + q;
        if (q + PD + 2 < kCflRows + 2) rr[q + PD + 2] = row(min(j + PD + 1, j1));
        if (q + PD < kCflRows) hh[q + PD] = halo(min(j + PD, j1 - 1));
        __builtin_amdgcn_sched_barrier(0);
        if (j < j1) {  // uniform over the workgroup (the last band may be shorter)
            const double2 pm = rr[q], pc = rr[q + 1], pp = rr[q + 2];
            const double hc = hh[q], a = a_node[j];
            // u = a d psi / d eta (rows j -+ 1), v = -d psi / dx (columns x -+ 1), as diag.cu:740-760
            const double l = wave_shr1(pc.y, hc), r = wave_shl1(pc.x, hc);  // columns x0+2l-1, x0+2l+2
            const double u0 = fabs(a * ((pp.x - pm.x) * hdy)), u1 = fabs(a * ((pp.y - pm.y) * hdy));
            const double v0 = fabs(-((pc.y - l) * hdx)), v1 = fabs(-((r - pc.x) * hdx));
            m0 = fmax(m0, fmax(u0, u1));
            m1 = fmax(m1, fmax(v0, v1));
            m2 = fmax(m2, fmax(v0 * a, v1 * a));
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    double m[3] = {m0, m1, m2};
    block_max<3>(m, sm);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) partial[(size_t)k * ntiles + b] = m[k];
}

__global__ void __launch_bounds__(kCflThreads) cfl_rows2_kernel(const double* __restrict__ psi, int Nx, int Ny,
                                                                double inv_dx, double inv_deta,
                                                                const double* __restrict__ a_node,
                                                                double* __restrict__ partial) {
    __shared__ double sm[3 * 16];
    cfl_tile_body(psi, Nx, Ny, inv_dx, inv_deta, a_node, xcd_tile(blockIdx.x, gridDim.x), gridDim.x, partial, sm);
}

// The same tiles for several fields (ensemble members, Nx a multiple of 512) in one launch:
// member m owns the global tiles [tile_begin, tile_begin + ntiles); the XCD renumbering runs over
// the global tile index. One workgroup per member then takes the maxima of its tiles.
struct CflTileJob {
    const double* psi;
    const double* a_node;
    double inv_dx, inv_deta;
    int Nx, Ny, tile_begin, ntiles;
    double* partial;  // [3][ntiles]
    double* out3;     // max|u|, max|v|, max|v| a
};
__global__ void __launch_bounds__(kCflThreads) cfl_tiles_group_kernel(const CflTileJob* __restrict__ jobs, int n) {
    __shared__ double sm[3 * 16];
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int m = member_of(b, n, [&](int k) { return jobs[k].tile_begin; });
    const CflTileJob& J = jobs[m];
    cfl_tile_body(J.psi, J.Nx, J.Ny, J.inv_dx, J.inv_deta, J.a_node, b - J.tile_begin, J.ntiles, J.partial, sm);
}
__global__ void __launch_bounds__(kCflThreads) cfl_tiles_final_group_kernel(const CflTileJob* __restrict__ jobs) {
    __shared__ double sm[3 * 16];
    const CflTileJob& J = jobs[blockIdx.x];
    double m[3];
    for (int k = 0; k < 3; ++k) {
        double v = 0.0;
        for (int i = threadIdx.x; i < J.ntiles; i += kCflThreads) v = fmax(v, J.partial[(size_t)k * J.ntiles + i]);
        m[k] = v;
    }
    block_max<3>(m, sm);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) J.out3[k] = m[k];
}

// One workgroup of 256 threads; each pass issues four loads per maximum before reducing them,
// so the n partials (1024 at 4096x2049) cost one memory latency instead of one per load
// (4.6 us per launch at 4096x2049; 5.5 us with 1024 threads).
constexpr int kCflFinalThreads = 256;
__global__ void __launch_bounds__(kCflFinalThreads) cfl_final_kernel(const double* __restrict__ partial, int n,
                                                                     double* __restrict__ scal, DtRule r) {
    __shared__ double sm[3 * 16];
    double m[3] = {0.0, 0.0, 0.0};
    for (int i0 = 0; i0 < n; i0 += 4 * kCflFinalThreads) {
        double v[3][4];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = i0 + u * kCflFinalThreads + (int)threadIdx.x;
                v[a][u] = (i < n) ? partial[(size_t)a * n + i] : 0.0;
            }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int u = 0; u < 4; ++u) m[a] = fmax(m[a], v[a][u]);
    }
    block_max<3>(m, sm);
    if (threadIdx.x == 0) {
        scal[2] = m[0];
        scal[3] = m[1];
        scal[4] = m[2];
        const double dt = dt_rule(m, r);
        scal[1] = dt;
        scal[0] = scal[0] + dt;
    }
}

// Ensembles: one thread per member; member m's scalars live at scal[8m]: t, dt, umax, vmax, vmaxm.
constexpr int kEnsScal = 8;
__global__ void ens_dt_kernel(const DtRule* __restrict__ rules, double* __restrict__ scal, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) scal[(size_t)m * kEnsScal + 1] = dt_rule(scal + (size_t)m * kEnsScal + 2, rules[m]);
}
__global__ void ens_advance_kernel(double* __restrict__ scal, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) scal[(size_t)m * kEnsScal] = scal[(size_t)m * kEnsScal] + scal[(size_t)m * kEnsScal + 1];
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
    int stage_tile = 10411;  // stage kernel shape (stage_tile_code; 0: generic kernel)
    // NSE2D_STAGE_FFT=1: production stages with the two-column kernel run fused with the forward
    // row transform (stage_fft_kernel) where the plan allows it. Off by default: at 4096x2049 the
    // fused kernel moves the fewest bytes (340 MB per launch against 246 + 134 MB for the pair)
    // but takes 84 us against 49 + 29 us, held to 3 wavefronts per SIMD by its 156 VGPRs.
    bool stage_fft = false;
    int prof_mask = 0;
    int prof_every = 1;         // time only every prof_every-th step (events cost time themselves)
    long long prof_steps = 0;   // steps since the last nse2d_channel_profile
    bool prof_now = true;       // this step is timed
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> ev_marks;  // (class, index of the start event)
    double prof_ms[NSE2D_PROF_NCLASS] = {0};
    long long prof_n[NSE2D_PROF_NCLASS] = {0};
    // optional closure term: before each stage closure_fn fills closure_buf from the stage state
    nse2d_closure_fn closure_fn = nullptr;
    void* closure_ctx = nullptr;
    float* closure_buf = nullptr;
    // per-channel row partials of the CFL maxima
    double* cfl_partial = nullptr;
    double* t_ptr() { return scalars + 0; }
    double* dt_ptr() { return scalars + 1; }
    double* max_ptr() { return scalars + 2; }
    double* cum_ptr() { return scalars + 5; }
    // wall closure of this channel (0: Thom, 1: free-slip): its descriptor's choice, else the
    // library-wide default at the time of the call
    int wall() const { return d.wall_type == 1 ? 0 : d.wall_type == 2 ? 1 : wall_type(); }
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
    ProfScope(ChannelImpl* ch, int k) : c(ch), cls(k), on(ch->prof_now && ((ch->prof_mask >> k) & 1)) {
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
    a.closure = c->closure_fn ? c->closure_buf : nullptr;
    if (d.mms) a.mp = make_mms_params(d.Nx, d.Lx, d.h, t_mms, d.nu, d.lin_drag, d.mms_A0, d.mms_eps, d.mms_Om, d.mms_kx);
    else a.mp = MmsParams{};
    return a;
}

// Shape of the wavefront stage kernel: RY rows per wavefront, SYW wavefronts per workgroup and
// prefetch depth PF, coded RY*100 + SYW*10 + PF, plus 10000 for the two-column kernels.
// NSE2D_STAGE_TILE=RYxSYWxPF (or RYxSYW, PF = 2; a leading 'w' for two columns per lane)
// selects another instantiated shape (used to tune on the device); 0 selects the generic kernel.
// Default: w4x1x1, one wavefront of 4 rows per workgroup (measured at 4096x2049 with the slabs
// swept in alternating directions: 132 us per step for the three stages against 139 us for
// w8x2x1, the earlier default, and 207 us for the best one-column shape 8x4x1).
// The manufactured-solution forcing (sin/cos per point) always runs in the generic kernel, so the
// production kernels carry none of its registers.
constexpr int kStageDefault = 841;        // one column per lane (ensembles, Nx % 128 != 0)
constexpr int kStageDefaultWide = 10411;  // two columns per lane, 4 rows x 1 wavefront
int stage_tile_code() {
    int ry = 4, sy = 1, pf = 1, wide = 1;
    if (const char* e = std::getenv("NSE2D_STAGE_TILE")) {
        if (e[0] == '0') return 0;
        wide = 0;  // one column per lane unless the shape says otherwise
        if (e[0] == 'w') { wide = 1; ++e; }
        else if (e[0] == 'p') { wide = 11; ++e; }  // ensembles: the persistent group kernel
        const int n = std::sscanf(e, "%dx%dx%d", &ry, &sy, &pf);
        if (n == 2) pf = 2;
        else if (n != 3) return kStageDefaultWide;
    }
    const int c = wide * 10000 + ry * 100 + sy * 10 + pf;
    switch (c) {
        case 841: case 842: case 843: case 1641: case 1642: case 1643: case 1621: case 1622: case 1623:
        case 3211: case 3212: case 3213:
        case 10841: case 10842: case 10441: case 11621: case 10821: case 10411: case 10421: case 10412:
        case 110821: case 110841: return c;
        default: return kStageDefaultWide;
    }
}

template <int STAGE, bool MMS>
void launch_stage_t(ChannelImpl* c, const StageArgs& a) {
    // the manufactured forcing and the closure term run in the generic kernel; the two-column
    // kernels (codes >= 10000) need Nx a multiple of 128, else the one-column kernel 8x4x1 runs
    int code = (!MMS && !a.closure && c->d.Nx % 64 == 0) ? c->stage_tile : 0;
    if (code >= 10000 && c->d.Nx % 128 != 0) code = kStageDefault;
#define NSE2D_STAGE_WAVE2(RYV, SYV, PFV)                                                                 \
    case 10000 + RYV * 100 + SYV * 10 + PFV:                                                             \
        stage_kernel_wave2<STAGE, RYV, SYV, PFV><<<(c->d.Nx / 128) * ceil_div(c->d.Ny, RYV * SYV),       \
                                                   dim3(64, SYV), 0, c->s>>>(a);                         \
        break;
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
        NSE2D_STAGE_WAVE2(8, 4, 1) NSE2D_STAGE_WAVE2(8, 4, 2) NSE2D_STAGE_WAVE2(4, 4, 1)
        NSE2D_STAGE_WAVE2(16, 2, 1) NSE2D_STAGE_WAVE2(8, 2, 1)
        NSE2D_STAGE_WAVE2(4, 1, 1) NSE2D_STAGE_WAVE2(4, 2, 1) NSE2D_STAGE_WAVE2(4, 1, 2)
        default: {
            const unsigned ntiles = ceil_div(c->d.Nx, SX) * ceil_div(c->d.Ny, SY * RY);
            stage_kernel<STAGE, MMS><<<ntiles, dim3(SX, SY), 0, c->s>>>(a);
        }
    }
#undef NSE2D_STAGE_WAVE
#undef NSE2D_STAGE_WAVE2
}

template <int STAGE>
void launch_stage(ChannelImpl* c, const StageArgs& a) {
    ProfScope ps(c, NSE2D_PROF_STAGE);
    if (a.mms) launch_stage_t<STAGE, true>(c, a);
    else launch_stage_t<STAGE, false>(c, a);
    NSE2D_LAUNCH_CHECK();
}

void close_stage(ChannelImpl* c, double* f, int stage, bool bookkeeping, bool forward_done) {
    const nse2d_channel_desc& d = c->d;
    if (!forward_done) { ProfScope ps(c, NSE2D_PROF_FFT_FWD); poisson_forward_interior(c->P, f, c->s); }
    { ProfScope ps(c, NSE2D_PROF_THOMAS); poisson_tridiag(c->P, d.psi_bot, d.psi_top, c->s); }
    if (!bookkeeping && poisson_fuses_closure(c->P)) {
        // the closure rides in the inverse transform's workgroups of rows 1 and Ny-2
        ProfScope ps(c, NSE2D_PROF_FFT_INV);
        poisson_inverse_closure(c->P, c->psi, f, d.d_dy_edge, d.psi_bot, d.psi_top, c->wall(), c->s);
        return;
    }
    { ProfScope ps(c, NSE2D_PROF_FFT_INV); poisson_inverse(c->P, c->psi, c->s); }
    ProfScope ps(c, NSE2D_PROF_THOM);
    if (bookkeeping) {
        thom_bookkeeping_kernel<<<1, 1024, 0, c->s>>>(f, c->psi, d.Nx, d.Ny, d.d_dy_edge, d.d_w_node, d.psi_bot,
                                                      d.psi_top, c->wall(), c->cum_ptr() + stage);
        NSE2D_LAUNCH_CHECK();
    } else {
        launch_thom(f, c->psi, d.Nx, d.Ny, d.d_dy_edge, d.psi_bot, d.psi_top, c->wall(), c->s);
    }
}

// The closure term of the stage state (omega_s, psi), when one is attached.
void closure_term(ChannelImpl* c, const double* ws) {
    if (c->closure_fn) c->closure_fn(c->closure_ctx, ws, c->psi, c->closure_buf, c->s);
}

void next_step_timed(ChannelImpl* c) { c->prof_now = (c->prof_steps++ % c->prof_every) == 0; }

bool stage_fft_plan(int M, int R) {
    switch (M * 100 + R) {
        case 256 * 100 + 4: case 512 * 100 + 8: case 1024 * 100 + 8: case 2048 * 100 + 8: case 4096 * 100 + 16:
        case 8192 * 100 + 16: return true;
        default: return false;
    }
}

// The fused stage + forward row transform when it applies (production stage, two-column stage
// shape, an instantiated row-FFT plan); false, having launched nothing, otherwise.
template <int STAGE>
bool launch_stage_fft(ChannelImpl* c, const StageArgs& a) {
    if (!c->stage_fft || c->stage_tile < 10000 || a.mms || a.closure || a.rhs_out) return false;
    int M = 0, R = 0;
    const void *twp = nullptr, *tw = nullptr;
    if (!poisson_row_fft(c->P, &M, &R, &twp, &tw) || !stage_fft_plan(M, R) || 2 * M != c->d.Nx) return false;
    auto* hat = static_cast<fft::cplx*>(poisson_hat(c->P));
    const auto* tp = static_cast<const fft::cplx*>(twp);
    const auto* tt = static_cast<const fft::cplx*>(tw);
    ProfScope ps(c, NSE2D_PROF_STAGE);
    switch (M * 100 + R) {
#define NSE2D_STAGE_FFT(MV, RV)                                                                         \
    case MV * 100 + RV:                                                                                 \
        stage_fft_kernel<STAGE, MV, RV><<<c->d.Ny, MV / RV, 0, c->s>>>(a, hat, tp, tt);                 \
        break;
        NSE2D_STAGE_FFT(256, 4) NSE2D_STAGE_FFT(512, 8) NSE2D_STAGE_FFT(1024, 8) NSE2D_STAGE_FFT(2048, 8)
        NSE2D_STAGE_FFT(4096, 16) NSE2D_STAGE_FFT(8192, 16)
#undef NSE2D_STAGE_FFT
    }
    NSE2D_LAUNCH_CHECK();
    return true;
}

template <int STAGE>
void stage_and_solve(ChannelImpl* c, const StageArgs& a, int stage, bool book) {
    const bool fused = launch_stage_fft<STAGE>(c, a);
    if (!fused) launch_stage<STAGE>(c, a);
    close_stage(c, a.out, stage, book, fused);
}

void do_step(ChannelImpl* c, double t, double dt, const double* dt_dev, int flags) {
    const bool book = (flags & NSE2D_STEP_WALL_BOOKKEEPING) != 0;
    double* rhs3 = (flags & NSE2D_STEP_STORE_RHS) ? c->rhs : nullptr;
    closure_term(c, c->w);
    stage_and_solve<1>(c, stage_args(c, c->w, nullptr, c->w1, nullptr, dt_dev, dt, t), 0, book);
    closure_term(c, c->w1);
    stage_and_solve<2>(c, stage_args(c, c->w1, c->w, c->w2, nullptr, dt_dev, dt, t + dt), 1, book);
    closure_term(c, c->w2);
    stage_and_solve<3>(c, stage_args(c, c->w2, c->w, c->w, rhs3, dt_dev, dt, t + 0.5 * dt), 2, book);
}

}  // namespace

}  // namespace nse2d

namespace nse2d {

// =============================================================================================
// Ensembles of independent channels (uneven sizes, own Re_b, IC and adaptive dt per member).
// One step = the CFL maxima of every member (one grouped launch), every member's dt (one launch),
// then per stage one grouped stage-kernel launch over all members and, per size class (members
// that share Nx, Ny, Lx, metrics and wall values, hence one Poisson plan), one grouped launch
// each of the forward transform, the Thomas solve and the inverse transform with the wall
// closure; finally every member's time advances. The sequence is captured once into a hipGraph
// and replayed. The grouped kernels run the single-channel kernels' bodies, so a member advances
// exactly as it would alone under the same kernel shapes.
// =============================================================================================
namespace {

struct EnsClass {
    int Nx = 0, Ny = 0;
    int wall = 0;  // 0: Thom closure, 1: free-slip (resolved when the ensemble is created)
    double Lx = 0, psi_bot = 0, psi_top = 0;
    const double *a_node = nullptr, *a_edge = nullptr, *dy_edge = nullptr;
    PoissonImpl* P = nullptr;
    std::vector<int> members;
    std::vector<void*> hats;
    void** d_hats = nullptr;   // [nc] spectrum buffers
    double** d_tab = nullptr;  // [4][nc]: omega, omega1, omega2, psi of the class's members
};

struct EnsMember {
    nse2d_channel_desc d;
    double *w = nullptr, *w1 = nullptr, *w2 = nullptr, *psi = nullptr, *f_row = nullptr;
    double dx = 0, deta = 0, inv_dx = 0, inv_deta = 0, inv_dx2 = 0, inv_deta2 = 0;
    int cls = 0;
    int wall = 0;
};

int stage_code_ry(int c) { return (c % 10000) / 100; }
int stage_code_sy(int c) { return (c / 10) % 10; }
bool stage_code_wide(int c) { return c >= 10000; }
bool stage_code_persistent(int c) { return c >= 100000; }  // ensembles: the persistent group kernel
std::string stage_code_name(int c) {
    return (stage_code_persistent(c) ? "p" : stage_code_wide(c) ? "w" : "") + std::to_string(stage_code_ry(c)) + "x" +
           std::to_string(stage_code_sy(c)) +
           "x" + std::to_string(c % 10);
}

}  // namespace

struct EnsembleImpl {
    hipStream_t s = nullptr;
    bool own_stream = false;
    std::vector<EnsMember> mem;
    std::vector<EnsClass> cls;
    double* scal = nullptr;        // [n][kEnsScal]: t, dt, umax, vmax, vmaxm
    DtRule* d_rules = nullptr;     // [n]
    StageArgs* d_args = nullptr;   // [3][n]
    // The members in size-class order: class k owns the slots [cls_off[k], cls_off[k+1]) of d_args
    // (per stage) and the first tiles d_tile_begin[cls_off[k] + k .. + nk] of its own launches.
    std::vector<int> cls_off;      // [ncls+1]
    std::vector<int> cls_tiles;    // [ncls]: tiles of one stage launch of class k
    int* d_tile_begin = nullptr;   // [n + ncls]
    // CFL maxima over 512 x kCflRows tiles of every member (members have Nx a power of two >= 512)
    CflTileJob* d_cfl_jobs = nullptr;  // [n]
    double* d_cfl_partial = nullptr;   // [3][tiles] per member
    int cfl_tiles = 0;                 // total tiles
    int stage_tile = kStageDefault;
    bool all_wide = false;  // every member has Nx % 128 == 0 (two-column stage kernels allowed)
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int cus = 256;  // compute units of the device (grid of the persistent stage kernel)
    // size classes after the first run their stages and Poisson solves on streams of their own
    // (default priority), forked from and joined to s by events
    std::vector<hipStream_t> cls_s;
    std::vector<hipEvent_t> cls_ev;  // [0]: the fork, [c]: class c done
};

namespace {

void ens_drop_graph(EnsembleImpl* e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

void ens_set_tiles(EnsembleImpl* e) {
    const int n = (int)e->mem.size();
    const int rows = stage_code_ry(e->stage_tile) * stage_code_sy(e->stage_tile);
    const int strip = stage_code_wide(e->stage_tile) ? 128 : 64;
    const int ncls = (int)e->cls.size();
    std::vector<int> tb;
    e->cls_tiles.assign(ncls, 0);
    for (int k = 0; k < ncls; ++k) {
        int t = 0;
        tb.push_back(0);
        for (int m : e->cls[k].members) {
            t += (e->mem[m].d.Nx / strip) * (int)ceil_div(e->mem[m].d.Ny, rows);
            tb.push_back(t);
        }
        e->cls_tiles[k] = t;
    }
    NSE2D_HIP_CHECK(hipMemcpyAsync(e->d_tile_begin, tb.data(), sizeof(int) * (n + ncls), hipMemcpyHostToDevice,
                                   e->s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(e->s));
}

// Stage `stage` of the members of size class k, on stream s.
void ens_launch_stage(EnsembleImpl* e, int stage, int k, hipStream_t s) {
    const int n = e->cls_off[k + 1] - e->cls_off[k];
    const StageArgs* args = e->d_args + (size_t)(stage - 1) * e->mem.size() + e->cls_off[k];
    const int* tile_begin = e->d_tile_begin + e->cls_off[k] + k;
    const int tiles = e->cls_tiles[k];
#define NSE2D_ENS_WAVE(RYV, SYV, PFV)                                                                      \
    case RYV * 100 + SYV * 10 + PFV:                                                                       \
        if (stage == 1)                                                                                    \
            stage_kernel_wave_group<1, RYV, SYV, PFV><<<tiles, dim3(64, SYV), 0, s>>>(args, tile_begin, n); \
        else if (stage == 2)                                                                               \
            stage_kernel_wave_group<2, RYV, SYV, PFV><<<tiles, dim3(64, SYV), 0, s>>>(args, tile_begin, n); \
        else                                                                                               \
            stage_kernel_wave_group<3, RYV, SYV, PFV><<<tiles, dim3(64, SYV), 0, s>>>(args, tile_begin, n); \
        break;
#define NSE2D_ENS_PERSIST(RYV, SYV, PFV)                                                                   \
    case 110000 + RYV * 100 + SYV * 10 + PFV: {                                                            \
        const int g = std::min(tiles, e->cus * (16 / SYV));                                       \
        if (stage == 1)                                                                                    \
            stage_kernel_wave2_group_persistent<1, RYV, SYV, PFV><<<g, dim3(64, SYV), 0, s>>>(args, tile_begin, n, tiles); \
        else if (stage == 2)                                                                               \
            stage_kernel_wave2_group_persistent<2, RYV, SYV, PFV><<<g, dim3(64, SYV), 0, s>>>(args, tile_begin, n, tiles); \
        else                                                                                               \
            stage_kernel_wave2_group_persistent<3, RYV, SYV, PFV><<<g, dim3(64, SYV), 0, s>>>(args, tile_begin, n, tiles); \
        break;                                                                                             \
    }
#define NSE2D_ENS_WAVE2(RYV, SYV, PFV)                                                                     \
    case 10000 + RYV * 100 + SYV * 10 + PFV:                                                               \
        if (stage == 1)                                                                                    \
            stage_kernel_wave2_group<1, RYV, SYV, PFV><<<tiles, dim3(64, SYV), 0, s>>>(args, tile_begin, n); \
        else if (stage == 2)                                                                               \
            stage_kernel_wave2_group<2, RYV, SYV, PFV><<<tiles, dim3(64, SYV), 0, s>>>(args, tile_begin, n); \
        else                                                                                               \
            stage_kernel_wave2_group<3, RYV, SYV, PFV><<<tiles, dim3(64, SYV), 0, s>>>(args, tile_begin, n); \
        break;
    switch (e->stage_tile) {
        NSE2D_ENS_WAVE(8, 4, 1) NSE2D_ENS_WAVE(8, 4, 2) NSE2D_ENS_WAVE(8, 4, 3)
        NSE2D_ENS_WAVE(16, 4, 1) NSE2D_ENS_WAVE(16, 4, 2) NSE2D_ENS_WAVE(16, 4, 3)
        NSE2D_ENS_WAVE(16, 2, 1) NSE2D_ENS_WAVE(16, 2, 2) NSE2D_ENS_WAVE(16, 2, 3)
        NSE2D_ENS_WAVE(32, 1, 1) NSE2D_ENS_WAVE(32, 1, 2) NSE2D_ENS_WAVE(32, 1, 3)
        NSE2D_ENS_WAVE2(8, 4, 1) NSE2D_ENS_WAVE2(8, 4, 2) NSE2D_ENS_WAVE2(4, 4, 1)
        NSE2D_ENS_WAVE2(16, 2, 1) NSE2D_ENS_WAVE2(8, 2, 1) NSE2D_ENS_WAVE2(4, 1, 1) NSE2D_ENS_WAVE2(4, 2, 1)
        NSE2D_ENS_PERSIST(8, 2, 1) NSE2D_ENS_PERSIST(8, 4, 1)
        default: std::abort();
    }
#undef NSE2D_ENS_WAVE
#undef NSE2D_ENS_WAVE2
#undef NSE2D_ENS_PERSIST
    NSE2D_LAUNCH_CHECK();
}

// Poisson solve and wall closure of the stage output of size class k's members, on stream s:
// field index 1 (omega1), 2 (omega2) or 0 (omega).
void ens_close(EnsembleImpl* e, int field, int k, hipStream_t s) {
    EnsClass& c = e->cls[k];
    const int nc = (int)c.members.size();
    double** tab = c.d_tab;
    poisson_forward_group(c.P, tab + (size_t)field * nc, c.d_hats, nc, s);
    poisson_tridiag_group(c.P, c.d_hats, nc, c.psi_bot, c.psi_top, s);
    poisson_inverse_group(c.P, c.d_hats, tab + (size_t)3 * nc, tab + (size_t)field * nc, nc, c.dy_edge, c.psi_bot,
                          c.psi_top, c.wall, s);
}

void ens_record_step(EnsembleImpl* e) {
    const int n = (int)e->mem.size();
    cfl_tiles_group_kernel<<<e->cfl_tiles, kCflThreads, 0, e->s>>>(e->d_cfl_jobs, n);
    NSE2D_LAUNCH_CHECK();
    cfl_tiles_final_group_kernel<<<n, kCflThreads, 0, e->s>>>(e->d_cfl_jobs);
    NSE2D_LAUNCH_CHECK();
    ens_dt_kernel<<<ceil_div(n, 64), 64, 0, e->s>>>(e->d_rules, e->scal, n);
    NSE2D_LAUNCH_CHECK();
    // The size classes share nothing between the dt and the time advance: each runs its three
    // stages and solves as a chain of its own, class k > 0 on stream cls_s[k], forked from and
    // joined to s by events (one graph branch per class when the step is captured), so that the
    // classes' launches overlap (1.664-1.681 -> 1.595-1.602 ms per step of the 64-member ensemble
    // on one host).
    const int ncls = (int)e->cls.size();
    if (ncls > 1) {
        NSE2D_HIP_CHECK(hipEventRecord(e->cls_ev[0], e->s));
        for (int k = 1; k < ncls; ++k) NSE2D_HIP_CHECK(hipStreamWaitEvent(e->cls_s[k], e->cls_ev[0], 0));
    }
    for (int k = 0; k < ncls; ++k) {
        hipStream_t s = k == 0 ? e->s : e->cls_s[k];
        ens_launch_stage(e, 1, k, s);
        ens_close(e, 1, k, s);
        ens_launch_stage(e, 2, k, s);
        ens_close(e, 2, k, s);
        ens_launch_stage(e, 3, k, s);
        ens_close(e, 0, k, s);
    }
    for (int k = 1; k < ncls; ++k) {
        NSE2D_HIP_CHECK(hipEventRecord(e->cls_ev[k], e->cls_s[k]));
        NSE2D_HIP_CHECK(hipStreamWaitEvent(e->s, e->cls_ev[k], 0));
    }
    ens_advance_kernel<<<ceil_div(n, 64), 64, 0, e->s>>>(e->scal, n);
    NSE2D_LAUNCH_CHECK();
}

void ens_steps(EnsembleImpl* e, int nsteps) {
    if (!e->use_graph) {
        for (int k = 0; k < nsteps; ++k) ens_record_step(e);
        return;
    }
    if (!e->exec) {
        NSE2D_HIP_CHECK(hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal));
        ens_record_step(e);
        NSE2D_HIP_CHECK(hipStreamEndCapture(e->s, &e->graph));
        NSE2D_HIP_CHECK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    }
    for (int k = 0; k < nsteps; ++k) NSE2D_HIP_CHECK(hipGraphLaunch(e->exec, e->s));
}

void ens_destroy(EnsembleImpl* e) {
    if (!e) return;
    if (e->s) (void)hipStreamSynchronize(e->s);
    ens_drop_graph(e);
    for (EnsMember& m : e->mem)
        for (double* f : {m.w, m.w1, m.w2, m.psi, m.f_row}) (void)hipFree(f);
    for (EnsClass& c : e->cls) {
        for (void* h : c.hats) (void)hipFree(h);
        (void)hipFree(c.d_hats);
        (void)hipFree(c.d_tab);
        poisson_destroy(c.P);
    }
    for (void* q : {(void*)e->scal, (void*)e->d_rules, (void*)e->d_args, (void*)e->d_tile_begin,
                    (void*)e->d_cfl_jobs, (void*)e->d_cfl_partial})
        (void)hipFree(q);
    for (hipStream_t q : e->cls_s)
        if (q) (void)hipStreamDestroy(q);
    for (hipEvent_t v : e->cls_ev)
        if (v) (void)hipEventDestroy(v);
    if (e->own_stream) (void)hipStreamDestroy(e->s);
    delete e;
}

EnsembleImpl* ens_create(const nse2d_channel_desc* descs, int n, hipStream_t s) {
    if (!descs || n < 1) { set_last_error("nse2d_ensemble_create: no members"); return nullptr; }
    for (int m = 0; m < n; ++m) {
        const nse2d_channel_desc& d = descs[m];
        if (d.Nx < 512 || d.Nx > 16384 || (d.Nx & (d.Nx - 1)) != 0 || d.Ny < 4) {
            set_last_error("nse2d_ensemble_create: member " + std::to_string(m) +
                           " needs Nx a power of two in [512, 16384] and Ny >= 4");
            return nullptr;
        }
        if (d.mms) { set_last_error("nse2d_ensemble_create: manufactured forcing needs host time"); return nullptr; }
        if (!d.d_a_node || !d.d_a_edge || !d.d_y_node || !d.d_dy_edge || !d.d_w_node) {
            set_last_error("nse2d_ensemble_create: member " + std::to_string(m) + " metric arrays must be set");
            return nullptr;
        }
        if (d.wall_type < 0 || d.wall_type > 2) {
            set_last_error("nse2d_ensemble_create: member " + std::to_string(m) + " wall_type must be 0, 1 or 2");
            return nullptr;
        }
    }
    auto* e = new EnsembleImpl;
    e->s = s;
    if (!s) {  // a captured graph needs a stream other than the null stream
        NSE2D_HIP_CHECK(hipStreamCreate(&e->s));
        e->own_stream = true;
    }
    e->mem.resize(n);
    for (int m = 0; m < n; ++m) {
        EnsMember& M = e->mem[m];
        M.d = descs[m];
        const nse2d_channel_desc& d = M.d;
        const int wall = d.wall_type == 1 ? 0 : d.wall_type == 2 ? 1 : wall_type();
        M.wall = wall;
        int k = 0;
        for (; k < (int)e->cls.size(); ++k) {
            const EnsClass& c = e->cls[k];
            if (c.Nx == d.Nx && c.Ny == d.Ny && c.Lx == d.Lx && c.a_node == d.d_a_node && c.a_edge == d.d_a_edge &&
                c.dy_edge == d.d_dy_edge && c.psi_bot == d.psi_bot && c.psi_top == d.psi_top && c.wall == wall)
                break;
        }
        if (k == (int)e->cls.size()) {
            EnsClass c;
            c.Nx = d.Nx; c.Ny = d.Ny; c.Lx = d.Lx; c.psi_bot = d.psi_bot; c.psi_top = d.psi_top; c.wall = wall;
            c.a_node = d.d_a_node; c.a_edge = d.d_a_edge; c.dy_edge = d.d_dy_edge;
            c.P = poisson_create_with(d.Nx, d.Ny, d.Lx, -1, -1, e->s);
            if (!c.P) { ens_destroy(e); return nullptr; }
            e->cls.push_back(c);
            poisson_enable_mapped(e->cls.back().P, d.d_a_node, d.d_a_edge, (double)(d.Ny - 1) / 2.0, e->s);
        }
        M.cls = k;
        e->cls[k].members.push_back(m);
        M.dx = d.Lx / (double)d.Nx;
        M.inv_dx = 1.0 / M.dx;
        M.inv_dx2 = M.inv_dx * M.inv_dx;
        M.deta = 2.0 / (double)(d.Ny - 1);
        M.inv_deta = 1.0 / M.deta;
        M.inv_deta2 = M.inv_deta * M.inv_deta;
        const size_t N = (size_t)d.Nx * d.Ny;
        for (double** f : {&M.w, &M.w1, &M.w2, &M.psi}) {
            NSE2D_HIP_CHECK(hipMalloc(f, N * sizeof(double)));
            NSE2D_HIP_CHECK(hipMemsetAsync(*f, 0, N * sizeof(double), e->s));
        }
        NSE2D_HIP_CHECK(hipMalloc(&M.f_row, d.Ny * sizeof(double)));
        forcing_rows_kernel<<<ceil_div(d.Ny, 256), 256, 0, e->s>>>(M.f_row, d.d_y_node, d.Ny, d.F0, d.nforce, d.h);
        NSE2D_LAUNCH_CHECK();
    }
    for (EnsClass& c : e->cls) {
        const int nc = (int)c.members.size();
        std::vector<double*> tab(4 * nc);
        for (int i = 0; i < nc; ++i) {
            void* h = nullptr;
            NSE2D_HIP_CHECK(hipMalloc(&h, poisson_hat_bytes(c.P)));
            c.hats.push_back(h);
            const EnsMember& M = e->mem[c.members[i]];
            tab[i] = M.w; tab[nc + i] = M.w1; tab[2 * nc + i] = M.w2; tab[3 * nc + i] = M.psi;
        }
        NSE2D_HIP_CHECK(hipMalloc(&c.d_hats, sizeof(void*) * nc));
        NSE2D_HIP_CHECK(hipMalloc(&c.d_tab, sizeof(double*) * 4 * nc));
        NSE2D_HIP_CHECK(hipMemcpyAsync(c.d_hats, c.hats.data(), sizeof(void*) * nc, hipMemcpyHostToDevice, e->s));
        NSE2D_HIP_CHECK(hipMemcpyAsync(c.d_tab, tab.data(), sizeof(double*) * 4 * nc, hipMemcpyHostToDevice, e->s));
        NSE2D_HIP_CHECK(hipStreamSynchronize(e->s));
    }
    NSE2D_HIP_CHECK(hipMalloc(&e->scal, sizeof(double) * kEnsScal * n));
    NSE2D_HIP_CHECK(hipMemsetAsync(e->scal, 0, sizeof(double) * kEnsScal * n, e->s));
    // the members' slots in size-class order
    std::vector<int> slot(n);
    e->cls_off.assign(1, 0);
    for (const EnsClass& c : e->cls) {
        for (size_t i = 0; i < c.members.size(); ++i) slot[c.members[i]] = e->cls_off.back() + (int)i;
        e->cls_off.push_back(e->cls_off.back() + (int)c.members.size());
    }
    e->cls_ev.assign(e->cls.size(), nullptr);
    e->cls_s.assign(e->cls.size(), nullptr);
    if (e->cls.size() > 1) {
        for (size_t k = 0; k < e->cls.size(); ++k)
            NSE2D_HIP_CHECK(hipEventCreateWithFlags(&e->cls_ev[k], hipEventDisableTiming));
        for (size_t k = 1; k < e->cls.size(); ++k)
            NSE2D_HIP_CHECK(hipStreamCreateWithFlags(&e->cls_s[k], hipStreamNonBlocking));
    }
    std::vector<DtRule> rules(n);
    std::vector<StageArgs> args(3 * (size_t)n);
    std::vector<CflTileJob> tj(n);
    int tiles = 0;
    for (int m = 0; m < n; ++m) {
        const EnsMember& M = e->mem[m];
        const nse2d_channel_desc& d = M.d;
        rules[m] = DtRule{M.dx, M.deta, d.cfl, d.cvisc, d.nu, M.inv_dx2, 1.0 / (d.dy_min * d.dy_min), d.dt_max};
        tj[m] = CflTileJob{M.psi, d.d_a_node, M.inv_dx, M.inv_deta, d.Nx, d.Ny, tiles,
                           (d.Nx / 512) * (int)ceil_div(d.Ny - 2, kCflRows), nullptr,
                           e->scal + (size_t)m * kEnsScal + 2};
        tiles += tj[m].ntiles;
        for (int st = 1; st <= 3; ++st) {
            StageArgs a{};
            a.psi = M.psi;
            a.ws = st == 1 ? M.w : (st == 2 ? M.w1 : M.w2);
            a.wn = st == 1 ? nullptr : M.w;
            a.out = st == 1 ? M.w1 : (st == 2 ? M.w2 : M.w);
            a.rhs_out = nullptr;
            a.Nx = d.Nx; a.Ny = d.Ny;
            a.inv_dx = M.inv_dx; a.inv_deta = M.inv_deta; a.inv_dx2 = M.inv_dx2; a.inv_deta2 = M.inv_deta2;
            a.nu = d.nu; a.lin_drag = d.lin_drag;
            a.a_node = d.d_a_node; a.a_edge = d.d_a_edge; a.f_row = M.f_row; a.y_node = d.d_y_node;
            a.dt_ptr = e->scal + (size_t)m * kEnsScal + 1;
            a.dt_val = 0.0;
            a.mms = 0;
            a.mp = MmsParams{};
            args[(size_t)(st - 1) * n + slot[m]] = a;
        }
    }
    NSE2D_HIP_CHECK(hipMalloc(&e->d_rules, sizeof(DtRule) * n));
    NSE2D_HIP_CHECK(hipMalloc(&e->d_args, sizeof(StageArgs) * 3 * n));
    NSE2D_HIP_CHECK(hipMalloc(&e->d_tile_begin, sizeof(int) * (n + e->cls.size())));
    NSE2D_HIP_CHECK(hipMemcpyAsync(e->d_rules, rules.data(), sizeof(DtRule) * n, hipMemcpyHostToDevice, e->s));
    NSE2D_HIP_CHECK(hipMemcpyAsync(e->d_args, args.data(), sizeof(StageArgs) * 3 * n, hipMemcpyHostToDevice, e->s));
    NSE2D_HIP_CHECK(hipMalloc(&e->d_cfl_partial, sizeof(double) * 3 * (size_t)tiles));
    for (int m = 0; m < n; ++m) tj[m].partial = e->d_cfl_partial + 3 * (size_t)tj[m].tile_begin;
    NSE2D_HIP_CHECK(hipMalloc(&e->d_cfl_jobs, sizeof(CflTileJob) * n));
    NSE2D_HIP_CHECK(hipMemcpyAsync(e->d_cfl_jobs, tj.data(), sizeof(CflTileJob) * n, hipMemcpyHostToDevice, e->s));
    e->cfl_tiles = tiles;
    e->all_wide = true;
    for (const EnsMember& M : e->mem) e->all_wide = e->all_wide && (M.d.Nx % 128 == 0);
    {
        int dev = 0, cus = 0;
        NSE2D_HIP_CHECK(hipGetDevice(&dev));
        NSE2D_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (cus > 0) e->cus = cus;
    }
    e->stage_tile = stage_tile_code();
    // the generic kernel has no group form; the two-column kernels need every Nx % 128 == 0
    if (e->stage_tile == 0 || (stage_code_wide(e->stage_tile) && !e->all_wide)) e->stage_tile = kStageDefault;
    ens_set_tiles(e);
    // the host tables above are copied asynchronously and go out of scope here
    NSE2D_HIP_CHECK(hipStreamSynchronize(e->s));
    return e;
}

void ens_init(EnsembleImpl* e) {
    for (EnsClass& c : e->cls) {
        const int nc = (int)c.members.size();
        poisson_forward_group(c.P, c.d_tab, c.d_hats, nc, e->s);
        poisson_tridiag_group(c.P, c.d_hats, nc, c.psi_bot, c.psi_top, e->s);
        poisson_inverse_group(c.P, c.d_hats, c.d_tab + (size_t)3 * nc, nullptr, nc, c.dy_edge, c.psi_bot, c.psi_top,
                              0, e->s);
    }
    for (EnsMember& M : e->mem)
        launch_thom(M.w, M.psi, M.d.Nx, M.d.Ny, M.d.d_dy_edge, M.d.psi_bot, M.d.psi_top, M.wall, e->s);
}

// Mean ms per step of nsteps un-captured steps after 2 warm-up steps.
double ens_time_steps(EnsembleImpl* e, int nsteps) {
    ens_drop_graph(e);
    const bool g = e->use_graph;
    e->use_graph = false;
    ens_steps(e, 2);
    hipEvent_t a, b;
    NSE2D_HIP_CHECK(hipEventCreate(&a));
    NSE2D_HIP_CHECK(hipEventCreate(&b));
    NSE2D_HIP_CHECK(hipEventRecord(a, e->s));
    ens_steps(e, nsteps);
    NSE2D_HIP_CHECK(hipEventRecord(b, e->s));
    NSE2D_HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    NSE2D_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e->use_graph = g;
    return ms / nsteps;
}

// Coordinate search over the stage-kernel shape, then each size class's (row-FFT radix, Thomas
// column width), timing whole ensemble steps. The state (omega, psi, t, dt) is saved before and
// restored after, so tuning does not advance the ensemble.
std::string ens_autotune(EnsembleImpl* e, int nsteps) {
    std::vector<std::vector<double>> save;
    for (EnsMember& M : e->mem) {
        const size_t N = (size_t)M.d.Nx * M.d.Ny;
        save.emplace_back(2 * N);
        NSE2D_HIP_CHECK(hipMemcpyAsync(save.back().data(), M.w, N * sizeof(double), hipMemcpyDeviceToHost, e->s));
        NSE2D_HIP_CHECK(hipMemcpyAsync(save.back().data() + N, M.psi, N * sizeof(double), hipMemcpyDeviceToHost, e->s));
    }
    std::vector<double> scal(kEnsScal * e->mem.size());
    NSE2D_HIP_CHECK(hipMemcpyAsync(scal.data(), e->scal, sizeof(double) * scal.size(), hipMemcpyDeviceToHost, e->s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(e->s));
    std::string rep;
    {
        const int codes[] = {841,   842,   843,   1641,  1642,  1643,  1621, 1622, 1623, 3211, 3212, 3213,
                             10821, 10841, 10842, 10441, 11621, 10411, 10421, 110821, 110841};
        int best = e->stage_tile;
        double tbest = 1e300;
        for (int code : codes) {
            if (stage_code_wide(code) && !e->all_wide) continue;
            e->stage_tile = code;
            ens_set_tiles(e);
            const double t = ens_time_steps(e, nsteps);
            if (t < tbest) { tbest = t; best = code; }
        }
        e->stage_tile = best;
        ens_set_tiles(e);
        rep += "stage " + stage_code_name(best);
    }
    for (EnsClass& c : e->cls) {
        int best_r = poisson_fft_radix(c.P), best_k = poisson_thomas_kc(c.P);
        double tbest = ens_time_steps(e, nsteps);
        for (int r : {4, 8, 16})
            for (int k : {8, 4, 2}) {
                PoissonImpl* P = poisson_create_with(c.Nx, c.Ny, c.Lx, r, k, e->s);
                if (!P) continue;
                if (poisson_fft_radix(P) != r || poisson_thomas_kc(P) != k || !poisson_group_supported(P)) {
                    poisson_destroy(P);
                    continue;
                }
                poisson_enable_mapped(P, c.a_node, c.a_edge, (double)(c.Ny - 1) / 2.0, e->s);
                std::swap(P, c.P);
                const double t = ens_time_steps(e, nsteps);
                if (t < tbest) { tbest = t; best_r = r; best_k = k; std::swap(P, c.P); }
                std::swap(P, c.P);
                NSE2D_HIP_CHECK(hipStreamSynchronize(e->s));
                poisson_destroy(P);
            }
        if (best_r != poisson_fft_radix(c.P) || best_k != poisson_thomas_kc(c.P)) {
            PoissonImpl* P = poisson_create_with(c.Nx, c.Ny, c.Lx, best_r, best_k, e->s);
            poisson_enable_mapped(P, c.a_node, c.a_edge, (double)(c.Ny - 1) / 2.0, e->s);
            NSE2D_HIP_CHECK(hipStreamSynchronize(e->s));
            poisson_destroy(c.P);
            c.P = P;
        }
        rep += "; " + std::to_string(c.Nx) + "x" + std::to_string(c.Ny) + ": fft radix " +
               std::to_string(poisson_fft_radix(c.P)) + ", thomas KC " + std::to_string(poisson_thomas_kc(c.P));
    }
    ens_drop_graph(e);
    for (size_t m = 0; m < e->mem.size(); ++m) {
        EnsMember& M = e->mem[m];
        const size_t N = (size_t)M.d.Nx * M.d.Ny;
        NSE2D_HIP_CHECK(hipMemcpyAsync(M.w, save[m].data(), N * sizeof(double), hipMemcpyHostToDevice, e->s));
        NSE2D_HIP_CHECK(hipMemcpyAsync(M.psi, save[m].data() + N, N * sizeof(double), hipMemcpyHostToDevice, e->s));
    }
    NSE2D_HIP_CHECK(hipMemcpyAsync(e->scal, scal.data(), sizeof(double) * scal.size(), hipMemcpyHostToDevice, e->s));
    NSE2D_HIP_CHECK(hipStreamSynchronize(e->s));
    return rep;
}

}  // namespace

}  // namespace nse2d

struct nse2d_channel {
    nse2d::ChannelImpl impl;
};
struct nse2d_ensemble {
    nse2d::EnsembleImpl* impl;
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
    if (d.wall_type < 0 || d.wall_type > 2) {
        set_last_error("nse2d_channel_create: wall_type must be 0 (library default), 1 (no-slip) or 2 (free-slip)");
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
    if (c->stage_tile >= 100000) c->stage_tile -= 100000;  // a single channel has no persistent form
    if (const char* e = std::getenv("NSE2D_STAGE_FFT")) c->stage_fft = (e[0] == '1');
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
    {   // one partial per row (cfl_rows_kernel) or per 512 x kCflRows tile (cfl_rows2_kernel)
        const size_t tiles = (d.Nx >= 512 ? (size_t)(d.Nx / 512) : 1) * ceil_div(d.Ny > 2 ? d.Ny - 2 : 1, kCflRows);
        NSE2D_HIP_CHECK(hipMalloc(&c->cfl_partial, 3 * std::max((size_t)d.Ny, tiles) * sizeof(double)));
    }
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
    (void)hipFree(c->closure_buf);
    (void)hipFree(c->cfl_partial);
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
    launch_thom(c->w, c->psi, d.Nx, d.Ny, d.d_dy_edge, d.psi_bot, d.psi_top, c->wall(), c->s);
}

void nse2d_channel_step(nse2d_channel_t ch, double t, double dt, int flags) {
    next_step_timed(&ch->impl);
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
    const DtRule rule{c->dx, c->deta, d.cfl, d.cvisc, d.nu, c->inv_dx2, inv_dymin2, d.dt_max};
    for (int k = 0; k < n; ++k) {
        next_step_timed(c);
        double* m = c->max_ptr();
        {
            ProfScope ps(c, NSE2D_PROF_CFL);
            if (d.Nx < 2 || d.Ny < 3) {  // diag.cu:826-832: zero maxima
                launch_max_uv(c->psi, d.Nx, d.Ny, c->inv_dx, c->inv_deta, d.d_a_node, m, m + 1, m + 2, c->s);
                dt_from_maxima_kernel<<<1, 64, 0, c->s>>>(m, c->dt_ptr(), rule);
                NSE2D_LAUNCH_CHECK();
                advance_time_kernel<<<1, 64, 0, c->s>>>(c->t_ptr(), c->dt_ptr());
            } else {
                // t advances before the stages: nothing in the step reads it (no manufactured forcing)
                int nparts = d.Ny;
                if (d.Nx % 512 == 0) {
                    nparts = (d.Nx / 512) * (int)ceil_div(d.Ny - 2, kCflRows);
                    cfl_rows2_kernel<<<nparts, kCflThreads, 0, c->s>>>(c->psi, d.Nx, d.Ny, c->inv_dx, c->inv_deta,
                                                                       d.d_a_node, c->cfl_partial);
                } else {
                    cfl_rows_kernel<<<d.Ny, kCflThreads, 0, c->s>>>(c->psi, d.Nx, d.Ny, c->inv_dx, c->inv_deta,
                                                                    d.d_a_node, c->cfl_partial);
                }
                NSE2D_LAUNCH_CHECK();
                cfl_final_kernel<<<1, kCflFinalThreads, 0, c->s>>>(c->cfl_partial, nparts, c->scalars, rule);
            }
            NSE2D_LAUNCH_CHECK();
        }
        do_step(c, 0.0, 0.0, c->dt_ptr(), 0);
    }
}

void nse2d_channel_profile(nse2d_channel_t ch, int class_mask) {
    ChannelImpl* c = &ch->impl;
    c->prof_mask = class_mask;
    c->prof_steps = 0;
    for (int k = 0; k < NSE2D_PROF_NCLASS; ++k) { c->prof_ms[k] = 0; c->prof_n[k] = 0; }
    c->ev_used = 0;
    c->ev_marks.clear();
}

void nse2d_channel_profile_every(nse2d_channel_t ch, int every) {
    ch->impl.prof_every = every > 1 ? every : 1;
    ch->impl.prof_steps = 0;
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

hipStream_t nse2d_channel_stream(nse2d_channel_t ch) { return ch->impl.s; }

int nse2d_channel_set_closure(nse2d_channel_t ch, nse2d_closure_fn fn, void* ctx) {
    ChannelImpl* c = &ch->impl;
    if (fn && !c->closure_buf) {
        const size_t N = (size_t)c->d.Nx * c->d.Ny;
        NSE2D_HIP_CHECK(hipMalloc(&c->closure_buf, N * sizeof(float)));
        NSE2D_HIP_CHECK(hipMemsetAsync(c->closure_buf, 0, N * sizeof(float), c->s));
    }
    c->closure_fn = fn;
    c->closure_ctx = fn ? ctx : nullptr;
    return 0;
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

nse2d_ensemble_t nse2d_ensemble_create(const nse2d_channel_desc* members, int n, hipStream_t s) {
    EnsembleImpl* e = ens_create(members, n, s);
    if (!e) return nullptr;
    return new nse2d_ensemble{e};
}
void nse2d_ensemble_destroy(nse2d_ensemble_t e) {
    if (!e) return;
    ens_destroy(e->impl);
    delete e;
}
int nse2d_ensemble_size(nse2d_ensemble_t e) { return (int)e->impl->mem.size(); }
double* nse2d_ensemble_field(nse2d_ensemble_t e, int member, int which) {
    if (member < 0 || member >= (int)e->impl->mem.size()) return nullptr;
    EnsMember& M = e->impl->mem[member];
    switch (which) {
        case NSE2D_FIELD_OMEGA: return M.w;
        case NSE2D_FIELD_PSI: return M.psi;
        case NSE2D_FIELD_W1: return M.w1;
        case NSE2D_FIELD_W2: return M.w2;
        default: return nullptr;
    }
}
double* nse2d_ensemble_scalars(nse2d_ensemble_t e) { return e->impl->scal; }
hipStream_t nse2d_ensemble_stream(nse2d_ensemble_t e) { return e->impl->s; }
void nse2d_ensemble_init(nse2d_ensemble_t e) { ens_init(e->impl); }
void nse2d_ensemble_step_auto(nse2d_ensemble_t e, int nsteps) { ens_steps(e->impl, nsteps); }
void nse2d_ensemble_set_graph(nse2d_ensemble_t e, int on) {
    ens_drop_graph(e->impl);
    e->impl->use_graph = on != 0;
}
void nse2d_ensemble_sync(nse2d_ensemble_t e) { NSE2D_HIP_CHECK(hipStreamSynchronize(e->impl->s)); }
int nse2d_ensemble_autotune(nse2d_ensemble_t e, int nsteps, char* report, size_t report_len) {
    const std::string r = ens_autotune(e->impl, nsteps < 1 ? 3 : nsteps);
    if (report && report_len) {
        std::snprintf(report, report_len, "%s", r.c_str());
    }
    return 0;
}

}  // extern "C"
