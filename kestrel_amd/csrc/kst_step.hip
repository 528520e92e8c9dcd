// kst_step.hip -- gfx950 kernels of the finite-volume step.
//
// One RK stage of the hydraulic operator is ONE kernel: a workgroup takes a
// 16x16 sub-block of an active tile, stages the stage state (plus a 2-cell
// halo strip on each side) in LDS, and fuses in one pass what the reference
// does in seven tile sweeps plus the stage update
// (src/HydraulicRHS.f90:64-136 and src/TimeStepper.f90:333-527):
//   limited slopes -> MUSCL faces -> CorrectSlopes -> desingularisation ->
//   central-upwind fluxes + CFL -> flux divergence + sources -> SSP-RK stage.
// Data in HBM is tile-major SoA: field f of cell (i,j) of tile t sits at
// f*FS + t*NC + j*nx + i, so a sub-block row is 128 contiguous bytes.
//
// Control (accept/reject, dt, tile activation) lives in device memory; the
// host enqueues whole steps and only reads back status every few steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kestrel_step.h"
#include "kst_physics.hpp"

#pragma clang fp contract(off)

namespace kst {

constexpr int B = 16;            // sub-block edge (cells)
constexpr int RW = B + 4;        // region width with 2-cell halo
constexpr int NREG = RW * RW;
constexpr int NTHREADS = 256;
constexpr int NFACE = (B + 1) * B;

enum Status { ST_NEW = 0, ST_RETRY = 1, ST_DONE = 2, ST_ERROR = 3 };

struct Ctrl {
    double t, tend, dt_retry, last_dt, min_dt, max_dt;
    unsigned long long slot[3];
    long long steps, rejected, rhs_evals, cell_updates;
    long long steps_left;
    int status, err, n_active, n_ghost, mode, pad;
};

struct Dev {
    double* S0;  // u^n, 4 fields
    double* S2;  // stage 2
    double* S3;  // stage 3
    double* R;   // RHS of u^n: E_w, E_rhu, E_rhv, E_hnpsi, I
    double *b0c, *btc, *cbx, *cby, *cgam;      // cell topography [tile][NC]
    double *xb0, *xbt, *xby, *xgam;            // x-face topography [tile][NFX]
    double *yb0, *ybt, *ybx, *ygam;            // y-face topography [tile][NFY]
    double *xnp, *xnm, *ynp, *ynm;             // CFL numerators gr^2*dx of the face's two cells
    double *pxu, *pxv, *pyv;                   // surface-gradient prefactors [tile][NC]
    double *vb0, *vbt;                         // tile vertex elevations [tile][NV]
    double *gsx, *gsy;                         // ghost w slopes [tile][NC]
    double* maxima;                            // [11][tile][NC]
    uint8_t* tstate;
    uint8_t* has_src;
    int* active;
    int* ghosts;
    unsigned* eflags;
    Ctrl* ctrl;
    // flux sources
    double *src_x, *src_y, *src_r, *src_t, *src_q, *src_p;
    int *src_n, *src_off, *src_cells;
    long long FS;  // field stride = ntiles*NC
};

// ---------------------------------------------------------------------------
// time-step control shared by kernels
// ---------------------------------------------------------------------------

// NextFluxSeriesTime (src/TimeStepper.f90:281-305).
__device__ double next_flux_time(const Dev& d, const Params& p, double t) {
    double nextT = kHuge, tdiff = kHuge;
    for (int i = 0; i < p.n_sources; ++i)
        for (int j = 0; j < d.src_n[i]; ++j) {
            double tt = d.src_t[d.src_off[i] + j];
            double tmp = tt - t;
            if (tmp > 0.0 && tmp < tdiff) { tdiff = tmp; nextT = tt; }
        }
    return nextT;
}

// ComputeAdvisedTimeStep (src/HydraulicRHS.f90:141-174) without the 0.9.
__device__ double advised_from_slot(const Params& p, unsigned long long slot, int n_active) {
    double M = bitsd(slot);
    if (n_active < p.ntiles) M = dmin(M, p.max_dt);
    double m = dmin(p.cfl * M, p.diff_ts);
    return dmin(m, p.max_dt);
}

// The step's dt: substep-1 advice clamped to the next output/flux-series time
// (src/TimeStepper.f90:155-168), or the refined dt of a retried step.
__device__ double step_dt(const Dev& d, const Params& p) {
    const Ctrl* c = d.ctrl;
    if (c->status == ST_RETRY) return c->dt_retry;
    double adv = 0.9 * advised_from_slot(p, c->slot[0], c->n_active);
    double tmax = dmin(c->tend, next_flux_time(d, p, c->t));
    return dmin(adv, tmax - c->t);
}

// ---------------------------------------------------------------------------
// explicit source terms (src/Equations.f90:456-620)
// ---------------------------------------------------------------------------
__device__ void flux_sources(const Dev& d, const Params& p, double t, double grid_t, double x,
                             double y, double& Qt, double& psiQt) {
    double qsum = 0.0, psum = 0.0;
    for (int J = 0; J < p.n_sources; ++J) {
        double Qf = 0.0, psiQf = 0.0;
        double sx = d.src_x[J], sy = d.src_y[J], sr = d.src_r[J];
        if (((x - sx) * (x - sx) + (y - sy) * (y - sy)) < sr * sr) {
            int n = d.src_n[J];
            const double* tm = d.src_t + d.src_off[J];
            const double* fl = d.src_q + d.src_off[J];
            const double* ps = d.src_p + d.src_off[J];
            if (n == 1) {
                if (t < tm[0] || (t == tm[0] && grid_t < t)) {
                    Qf = 0.0; psiQf = 0.0;
                } else {
                    Qf = fl[0];
                    psiQf = ps[0] * Qf;
                }
            } else {
                if (t < tm[0] || t > tm[n - 1] || (t == tm[0] && grid_t < t) ||
                    (t == tm[n - 1] && grid_t == t)) {
                    Qf = 0.0; psiQf = 0.0;
                } else {
                    for (int K = 1; K < n; ++K)
                        if (t >= tm[K - 1] && t <= tm[K]) {
                            double Qa = fl[K - 1], pa = ps[K - 1], ta = tm[K - 1];
                            double Qb = fl[K], pb = ps[K], tb = tm[K];
                            Qf = Qa + (Qb - Qa) * (t - ta) / (tb - ta);
                            double pf = pa + (pb - pa) * (t - ta) / (tb - ta);
                            psiQf = pf * Qf;
                        }
                }
            }
        }
        Qf = Qf / (double)d.src_cells[J] / p.dx / p.dy;
        psiQf = psiQf / (double)d.src_cells[J] / p.dx / p.dy;
        qsum = qsum + Qf;
        psum = psum + psiQf;
    }
    Qt = 0.0 + qsum;
    psiQt = 0.0 + psum;
}

// ---------------------------------------------------------------------------
// the fused stage kernel
// ---------------------------------------------------------------------------

// Limiters at compile time (src/Limiters.f90:83-184).  MinMod1/2 use the
// identity max(-x,-y) = -min(x,y): negation and |.| are exact, so the value
// equals the reference's branchy form.
template <int LIM>
__device__ __forceinline__ double lim(double a, double b) {
    if (LIM == KST_LIMITER_MINMOD2) {
        const double th = 1.3;
        const double m = fmin(th * fabs(a), fmin(th * fabs(b), 0.5 * fabs(a + b)));
        return (a * b <= 0.0) ? 0.0 : copysign(m, a);
    } else if (LIM == KST_LIMITER_MINMOD1) {
        const double m = fmin(fabs(a), fabs(b));
        return (a * b <= 0.0) ? 0.0 : copysign(m, a);
    } else {
        return limiter(LIM, a, b);
    }
}

template <int NQ>
struct SmemT {
    double w[NREG], h[NREG], u[NREG], v[NREG], rho[NREG], gc[NREG];
    double xbed[B][B + 3];   // bed of x-faces -1..B+1 of each row (relative)
    double ybed[B][B + 3];   // bed of y-faces -1..B+1 of each column
    int gidx[NREG];
    uint8_t act[NREG];
    double fx[NQ][NFACE];    // hw, hru, hrv, hh, g (, pu, pv with eddy viscosity)
    double fy[NQ][NFACE];
    double red[NTHREADS / 64];
    int tile, bx0, by0, bxa, bya, skip;
    double dt, tstage;
};

__device__ __forceinline__ int rg(int rx, int ry) { return (ry + 2) * RW + (rx + 2); }

// Map a cell given relative to tile t (i in [-2, nx+1], j in [-2, ny+1], not
// both outside) to (tile, ci, cj).
__device__ __forceinline__ void map_cell(const Params& p, int t, int i, int j, int& tile, int& ci,
                                         int& cj) {
    tile = t; ci = i; cj = j;
    if (i < 0) { tile -= 1; ci += p.nx; }
    else if (i >= p.nx) { tile += 1; ci -= p.nx; }
    if (j < 0) { tile -= p.ntx; cj += p.ny; }
    else if (j >= p.ny) { tile += p.ntx; cj -= p.ny; }
}

// Fused work of one x- or y-face (CalculateLimitedDerivs + Reconstruct +
// CorrectSlopes + CalculateFluxes, src/HydraulicRHS.f90:181-1182).
// Along the face normal the four cells are c[-2], L=c[-1], R=c[0], c[1];
// fi is the face position along the normal, fj the row (x) / column (y).
template <bool YDIR, bool VISC, int LIM, bool CFL, typename Smem>
__device__ __forceinline__ void face_flux(const Dev& d, const Params& p, Smem& s, int fi, int fj,
                                          double& cfl_min) {
    int iLL, iL, iR, iRR;
    if (!YDIR) { iLL = rg(fi - 2, fj); iL = rg(fi - 1, fj); iR = rg(fi, fj); iRR = rg(fi + 1, fj); }
    else       { iLL = rg(fj, fi - 2); iL = rg(fj, fi - 1); iR = rg(fj, fi); iRR = rg(fj, fi + 1); }
    const double dxr = YDIR ? p.dyr : p.dxr;
    const double dx = YDIR ? p.dy : p.dx;
    const bool actL = s.act[iL], actR = s.act[iR];
    const int nn = YDIR ? p.ny : p.nx;
    const int fpos = (YDIR ? s.by0 : s.bx0) + fi;     // face index within the tile, 0..nn

    // --- slopes (CalculateLimitedDerivs, src/HydraulicRHS.f90:181-386);
    // non-active tiles keep their stored w slope and zero slopes otherwise.
    const double wLL = s.w[iLL], wL = s.w[iL], wR = s.w[iR], wRR = s.w[iRR];
    const double* gs = YDIR ? d.gsy : d.gsx;
    const double sLw = actL ? dxr * lim<LIM>(wR - wL, wL - wLL) : gs[s.gidx[iL]];
    const double sRw = actR ? dxr * lim<LIM>(wRR - wR, wR - wL) : gs[s.gidx[iR]];
    const double hLL = s.h[iLL], hL = s.h[iL], hR = s.h[iR], hRR = s.h[iRR];
    const double sLh = actL ? dxr * lim<LIM>(hR - hL, hL - hLL) : 0.0;
    const double sRh = actR ? dxr * lim<LIM>(hRR - hR, hR - hL) : 0.0;
    const double uLL = s.u[iLL], uL = s.u[iL], uR = s.u[iR], uRR = s.u[iRR];
    const double sLu = actL ? dxr * lim<LIM>(uR - uL, uL - uLL) : 0.0;
    const double sRu = actR ? dxr * lim<LIM>(uRR - uR, uR - uL) : 0.0;
    const double vLL = s.v[iLL], vL = s.v[iL], vR = s.v[iR], vRR = s.v[iRR];
    const double sLv = actL ? dxr * lim<LIM>(vR - vL, vL - vLL) : 0.0;
    const double sRv = actR ? dxr * lim<LIM>(vRR - vR, vR - vL) : 0.0;
    const double rLL = s.rho[iLL], rL = s.rho[iL], rR = s.rho[iR], rRR = s.rho[iRR];
    const double sLr = actL ? dxr * lim<LIM>(rR - rL, rL - rLL) : 0.0;
    const double sRr = actR ? dxr * lim<LIM>(rRR - rR, rR - rL) : 0.0;

    // --- MUSCL face values (Reconstruct, src/HydraulicRHS.f90:397-490)
    const double wEL = wL + sLw * 0.5 * dx, wWL = wL - sLw * 0.5 * dx;   // L's east / west
    const double wWR = wR - sRw * 0.5 * dx, wER = wR + sRw * 0.5 * dx;   // R's west / east

    // beds of the faces of L and R (InterpolateB == face b0 when bt = 0,
    // src/HydraulicRHS.f90:741-753); the face itself is entry fi + 1
    const double* bed = YDIR ? &s.ybed[fj][0] : &s.xbed[fj][0];
    const double bWL = bed[fi], fb0 = bed[fi + 1], bER = bed[fi + 2];

    // --- CorrectSlopes for w (src/HydraulicRHS.f90:560-736): each cell's
    // pair of face values is replaced by the bed slope if either is below bed.
    double wFL = wEL, wFR = wWR;
    if ((wEL < fb0) || (wWL < bWL)) wFL = wL + 0.5 * (fb0 - bWL);
    if ((wER < bER) || (wWR < fb0)) wFR = wR + 0.5 * (fb0 - bER);
    // Hnpsi: inside the tile both faces of a cell are reset; across the
    // tile edge only the copy this tile holds is checked.
    double hFL = hL + sLh * 0.5 * dx, hFR = hR - sRh * 0.5 * dx;
    if (fpos == 0) {
        if (hFL < 0.0) hFL = hL;
    } else {
        const double hWL = hL - sLh * 0.5 * dx;
        if ((hFL < 0.0) || (hWL < 0.0)) hFL = hL;
    }
    if (fpos == nn) {
        if (hFR < 0.0) hFR = hR;
    } else {
        const double hER = hR + sRh * 0.5 * dx;
        if ((hFR < 0.0) || (hER < 0.0)) hFR = hR;
    }
    const double uFL = uL + sLu * 0.5 * dx, uFR = uR - sRu * 0.5 * dx;
    const double vFL = vL + sLv * 0.5 * dx, vFR = vR - sRv * 0.5 * dx;
    const double rFL = rL + sLr * 0.5 * dx, rFR = rR - sRr * 0.5 * dx;

    // face topography of this tile
    const int t = s.tile;
    int fidx;
    double fbp, fg;
    if (!YDIR) {
        fidx = t * p.NFX + (s.by0 + fj) * (p.nx + 1) + fpos;
        fbp = d.xby[fidx]; fg = d.xgam[fidx];
    } else {
        fidx = t * p.NFY + fpos * p.nx + (s.bx0 + fj);
        fbp = d.ybx[fidx]; fg = d.ygam[fidx];
    }
    const double fbt = 0.0;
    // face Hn from the corrected w (src/HydraulicRHS.f90:492-517)
    const double HnL = compute_hn(wFL, fb0, fbt, fg);
    const double HnR = compute_hn(wFR, fb0, fbt, fg);
    const double unL = YDIR ? vFL : uFL, unR = YDIR ? vFR : uFR;
    const double ruL = rFL * HnL * uFL, ruR = rFR * HnR * uFR;
    const double rvL = rFL * HnL * vFL, rvR = rFR * HnR * vFR;

    // --- wave speeds and CFL (src/HydraulicRHS.f90:949-1006)
    const double rP = wave_root(p, HnR, fbp, fg);
    const double rM = wave_root(p, HnL, fbp, fg);
    double wsP = unR + rP, wsM = unL + rM;
    double aPos = (wsP > wsM) ? wsP : wsM;
    if (aPos < 0.0) aPos = 0.0;
    wsP = unR - rP; wsM = unL - rM;
    double aNeg = (wsP < wsM) ? wsP : wsM;
    if (aNeg > 0.0) aNeg = 0.0;
    if (CFL) {
        // numerators (min(gam_cell/gam_face,1))^2 * dx are static topography
        if (aPos > kEps) cfl_min = fmin((YDIR ? d.ynp : d.xnp)[fidx] / aPos, cfl_min);
        if (fabs(aNeg) > kEps) cfl_min = fmin((YDIR ? d.ynm : d.xnm)[fidx] / fabs(aNeg), cfl_min);
    }

    // --- fluxes (src/HydraulicRHS.f90:1027-1067, src/Equations.f90:53-245)
    double* F = YDIR ? &s.fy[0][0] : &s.fx[0][0];
    const int fslot = YDIR ? (fi * B + fj) : (fj * (B + 1) + fi);
    const double dif = aPos - aNeg;
    if (dif < 1e-10) {
        for (int q = 0; q < (VISC ? 7 : 5); ++q) F[q * NFACE + fslot] = 0.0;
        return;
    }
    const double cPw = HnR * unR * fg, cMw = HnL * unL * fg;
    const double cPh = hFR * unR * fg, cMh = hFL * unL * fg;
    const double cPu = ruR * unR, cMu = ruL * unL;
    const double cPv = rvR * unR, cMv = rvL * unL;
    double hp = -fbt;
    hp = hp + (wFR - fb0);
    const double hydP = 0.5 * p.g * rFR * hp * hp;
    hp = -fbt;
    hp = hp + (wFL - fb0);
    const double hydM = 0.5 * p.g * rFL * hp * hp;

    double x = HnR * fg - HnL * fg;
    x = x * aPos * aNeg;
    x = x + (aPos * cMw - aNeg * cPw);
    F[0 * NFACE + fslot] = x / dif;
    x = ruR - ruL;
    x = x * aPos * aNeg;
    x = x + (aPos * cMu - aNeg * cPu);
    F[1 * NFACE + fslot] = x / dif;
    x = rvR - rvL;
    x = x * aPos * aNeg;
    x = x + (aPos * cMv - aNeg * cPv);
    F[2 * NFACE + fslot] = x / dif;
    x = hFR * fg - hFL * fg;
    x = x * aPos * aNeg;
    x = x + (aPos * cMh - aNeg * cPh);
    F[3 * NFACE + fslot] = x / dif;
    F[4 * NFACE + fslot] = (aPos * hydM - aNeg * hydP) / dif;
    if (VISC) {
        double pPu, pPv, pMu, pMv;
        if (HnR < 0.0) { pPu = 0.0; pPv = 0.0; }
        else { pPu = p.nu * rFR * HnR * sRu; pPv = p.nu * rFR * HnR * sRv; }
        if (HnL < 0.0) { pMu = 0.0; pMv = 0.0; }
        else { pMu = p.nu * rFL * HnL * sLu; pMv = p.nu * rFL * HnL * sLv; }
        F[5 * NFACE + fslot] = 0.5 * (pPu + pMu);
        F[6 * NFACE + fslot] = 0.5 * (pPv + pMv);
    }
}

// Block-wide minimum of a positive double into a control slot.
template <typename Smem>
__device__ void reduce_min_to(Smem& s, double v, unsigned long long* slot) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) s.red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = s.red[0];
        for (int i = 1; i < NTHREADS / 64; ++i) m = fmin(m, s.red[i]);
        if (m < kHuge) atomicMin(slot, dbits(m));
    }
}

// Bed of a face given by its tile-relative index K in [-1, nn+1] along the
// normal and its row/column r (single-valued b0: a tile's edge faces equal
// its neighbours').
template <bool YDIR>
__device__ __forceinline__ double face_bed(const Dev& d, const Params& p, int t, int K, int r) {
    if (!YDIR) {
        if (K < 0) return d.xb0[(t - 1) * p.NFX + r * (p.nx + 1) + (p.nx - 1)];
        if (K > p.nx) return d.xb0[(t + 1) * p.NFX + r * (p.nx + 1) + 1];
        return d.xb0[t * p.NFX + r * (p.nx + 1) + K];
    } else {
        if (K < 0) return d.yb0[(t - p.ntx) * p.NFY + (p.ny - 1) * p.nx + r];
        if (K > p.ny) return d.yb0[(t + p.ntx) * p.NFY + 1 * p.nx + r];
        return d.yb0[t * p.NFY + K * p.nx + r];
    }
}

// STAGE 1: RHS(u^n) -> R.            (IntegrateTo, src/TimeStepper.f90:155)
// STAGE 2: u1 = f(u^n, R); RHS(u1); u2 -> S2.   (src/TimeStepper.f90:371-444)
// STAGE 3: RHS(u2); u3 -> S3.                   (src/TimeStepper.f90:447-498)
// STAGE 4: RHS(u3); implicit drag -> S0; maxima; activation flags.
//                                               (src/TimeStepper.f90:501-524)
template <int STAGE, bool VISC, int LIM>
__global__ __launch_bounds__(NTHREADS) void stage_kernel(Dev d, Params p) {
    __shared__ SmemT<VISC ? 7 : 5> s;
    const int tid = threadIdx.x;
    if (tid == 0) {
        const Ctrl* c = d.ctrl;
        int skip = 0;
        if (c->status == ST_DONE || c->status == ST_ERROR) skip = 1;
        if (STAGE == 1 && c->status != ST_NEW) skip = 1;
        double dt = 0.0;
        if (!skip && STAGE >= 2) {
            dt = step_dt(d, p);
            if (STAGE >= 3 && advised_from_slot(p, c->slot[1], c->n_active) < dt) skip = 1;
            if (STAGE >= 4 && advised_from_slot(p, c->slot[2], c->n_active) < dt) skip = 1;
        }
        s.dt = dt;
        s.tstage = (STAGE == 1) ? c->t : (STAGE == 3 ? c->t + 0.5 * dt : c->t + dt);
        // work item -> (tile, sub-block); contiguous chunks of the active
        // list go to the same XCD (blocks are dealt round-robin to 8 XCDs)
        const int sbx = (p.nx + B - 1) / B, sby = (p.ny + B - 1) / B, spt = sbx * sby;
        const int nwork = c->n_active * spt;
        const int per = (nwork + 7) / 8;
        const int b = blockIdx.x;
        const int w = (b % 8) * per + b / 8;
        if (b / 8 >= per || w >= nwork) skip = 1;
        if (!skip) {
            s.tile = d.active[w / spt];
            const int sub = w % spt;
            s.bx0 = (sub % sbx) * B;
            s.by0 = (sub / sbx) * B;
            s.bxa = min(B, p.nx - s.bx0);
            s.bya = min(B, p.ny - s.by0);
        }
        s.skip = skip;
    }
    __syncthreads();
    if (s.skip) return;
    const int t = s.tile;
    const double dt = s.dt;
    const int FS = (int)d.FS;
    const int bxa = s.bxa, bya = s.bya;

    // ---- phase A: stage state + derived variables on the region ----------
    const int tx = tid % B, ty = tid / B;
    const bool own = (tx < bxa) && (ty < bya);
    double ow = 0, oru = 0, orv = 0, oh = 0, ohn = 0, ou = 0, ov = 0, opsi = 0, orho = 0;
    double ob0 = 0, og = 1;
    double pw = 0, pru = 0, prv = 0, ph = 0;   // u^n of the owned cell (stage 2)
    int oidx = 0;
    for (int round = 0; round < 2; ++round) {
        int rx, ry;
        bool valid;
        if (round == 0) {
            rx = tx; ry = ty; valid = own;
        } else {
            if (tid >= 128) break;
            const int strip = tid >> 5, k = tid & 31, dd = k >> 4, pos = k & 15;
            if (strip == 0) { rx = -2 + dd; ry = pos; valid = pos < bya; }
            else if (strip == 1) { rx = bxa + dd; ry = pos; valid = pos < bya; }
            else if (strip == 2) { rx = pos; ry = -2 + dd; valid = pos < bxa; }
            else { rx = pos; ry = bya + dd; valid = pos < bxa; }
        }
        if (!valid) continue;
        int tc, ci, cj;
        map_cell(p, t, s.bx0 + rx, s.by0 + ry, tc, ci, cj);
        const int gi = tc * p.NC + cj * p.nx + ci;
        double w, ru, rv, h;
        if (STAGE == 1 || STAGE == 2) {
            w = d.S0[gi]; ru = d.S0[FS + gi]; rv = d.S0[2 * FS + gi]; h = d.S0[3 * FS + gi];
            if (STAGE == 2) {
                if (round == 0) { pw = w; pru = ru; prv = rv; ph = h; }
                const double ew = d.R[gi], eu = d.R[FS + gi], ev = d.R[2 * FS + gi];
                const double eh = d.R[3 * FS + gi], im = d.R[4 * FS + gi];
                ru = (ru + dt * eu) / (1.0 - dt * im);
                rv = (rv + dt * ev) / (1.0 - dt * im);
                w = w + dt * ew;
                h = h + dt * eh;
            }
        } else {
            const double* S = (STAGE == 3) ? d.S2 : d.S3;
            w = S[gi]; ru = S[FS + gi]; rv = S[2 * FS + gi]; h = S[3 * FS + gi];
        }
        const double b0c = d.b0c[gi], gam = d.cgam[gi];
        // ComputeDesingularisedVariables (src/HydraulicRHS.f90:762-878)
        double hn = compute_hn(w, b0c, 0.0, gam);
        if (hn < 0.0) hn = 0.0;
        double hpc = h;
        if (hpc < 0.0) hpc = 0.0;
        const double eps = p.hthr;
        const double den = hn * hn + dmax(hn * hn, eps * eps);
        const double psi = dmin(2.0 * hn * hpc / den, p.max_pack);
        const double rho = density(p, psi);
        const double uu = 2.0 * hn * ru / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
        const double vv = 2.0 * hn * rv / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
        const int r = rg(rx, ry);
        s.w[r] = w; s.h[r] = h; s.u[r] = uu; s.v[r] = vv; s.rho[r] = rho; s.gc[r] = gam;
        s.gidx[r] = gi;
        s.act[r] = (d.tstate[tc] == KST_TILE_ACTIVE) ? 1 : 0;
        if (round == 0) {
            ow = w; oru = ru; orv = rv; oh = h; ohn = hn; ou = uu; ov = vv; opsi = psi; orho = rho;
            ob0 = b0c; og = gam; oidx = gi;
        }
    }
    // face beds: rows of x-faces -1..bxa+1 and columns of y-faces -1..bya+1
    for (int q = tid; q < B * (B + 3); q += NTHREADS) {
        const int r = q / (B + 3), k = q % (B + 3) - 1;
        if (r < bya && k <= bxa + 1) s.xbed[r][k + 1] = face_bed<false>(d, p, t, s.bx0 + k, s.by0 + r);
        if (r < bxa && k <= bya + 1) s.ybed[r][k + 1] = face_bed<true>(d, p, t, s.by0 + k, s.bx0 + r);
    }
    __syncthreads();

    // ---- phase B: face fluxes --------------------------------------------
    double cflm = kHuge;
    // x-faces: consecutive lanes walk along a row; y-faces: along a row too
    if (bxa == B && bya == B) {
        for (int f = tid; f < (B + 1) * B; f += NTHREADS)
            face_flux<false, VISC, LIM, STAGE != 4>(d, p, s, f % (B + 1), f / (B + 1), cflm);
        for (int f = tid; f < (B + 1) * B; f += NTHREADS)
            face_flux<true, VISC, LIM, STAGE != 4>(d, p, s, f / B, f % B, cflm);
    } else {
        for (int f = tid; f < (bxa + 1) * bya; f += NTHREADS)
            face_flux<false, VISC, LIM, STAGE != 4>(d, p, s, f % (bxa + 1), f / (bxa + 1), cflm);
        for (int f = tid; f < (bya + 1) * bxa; f += NTHREADS)
            face_flux<true, VISC, LIM, STAGE != 4>(d, p, s, f / bxa, f % bxa, cflm);
    }
    if (STAGE != 4) {
        reduce_min_to(s, cflm, &d.ctrl->slot[STAGE - 1]);
    } else {
        __syncthreads();
    }

    // ---- phase C: RHS of the owned cell and the stage update -------------
    if (!own) return;
    const int xs = ty * (B + 1) + tx, ys = ty * B + tx;
    const double* FXp = &s.fx[0][0];
    const double* FYp = &s.fy[0][0];
#define FX(q, o) FXp[(q) * NFACE + xs + (o)]
#define FY(q, o) FYp[(q) * NFACE + ys + (o)]
    const double gam = og;
    const double bx = d.cbx[oidx], by = d.cby[oidx];
    // prefactors (1+by^2)/gam, -(bx by/gam), (1+bx^2)/gam are static topography
    const double gxu = d.pxu[oidx], gxv = d.pxv[oidx], gyu = gxv, gyv = d.pyv[oidx];
    // ConstructHydraulicRHS (src/HydraulicRHS.f90:1228-1271)
    const double stw = (FX(0, 0) - FX(0, 1)) * p.dxr / (gam * gam) + (FY(0, 0) - FY(0, B)) * p.dyr / (gam * gam);
    const double sth = (FX(3, 0) - FX(3, 1)) * p.dxr / gam + (FY(3, 0) - FY(3, B)) * p.dyr / gam;
    double stu, stv;
    {
        Kahan k1;
        k1.add(FX(1, 0) - FX(1, 1));
        k1.add((FX(4, 0) - FX(4, 1)) * gxu);
        k1.add(VISC ? (FX(5, 1) - FX(5, 0)) : 0.0);
        stu = k1.s * p.dxr;
        Kahan k2;
        k2.add(FY(1, 0) - FY(1, B));
        k2.add((FY(4, 0) - FY(4, B)) * gyu);
        k2.add(VISC ? (FY(5, B) - FY(5, 0)) : 0.0);
        stu = stu + k2.s * p.dyr;
        Kahan k3;
        k3.add(FX(2, 0) - FX(2, 1));
        k3.add((FX(4, 0) - FX(4, 1)) * gxv);
        k3.add(VISC ? (FX(6, 1) - FX(6, 0)) : 0.0);
        stv = k3.s * p.dxr;
        Kahan k4;
        k4.add(FY(2, 0) - FY(2, B));
        k4.add((FY(4, 0) - FY(4, B)) * gyv);
        k4.add(VISC ? (FY(6, B) - FY(6, 0)) : 0.0);
        stv = stv + k4.s * p.dyr;
    }
#undef FX
#undef FY
    // ExplicitSourceTerms (src/Equations.f90:456-620); with no source the
    // flux terms are 0/(gam*gam) = 0 exactly
    double Ew, Eh;
    if (p.n_sources > 0 && d.has_src[t]) {
        double Qt = 0.0, psiQt = 0.0;
        const int ti = t % p.ntx, tj = t / p.ntx;
        const int i = s.bx0 + tx, j = s.by0 + ty;
        const double x = -0.5 * p.x_size + p.dx * (((ti + 1) - 1.0) * p.nx + ((i + 1) - 0.5));
        const double y = -0.5 * p.y_size + p.dy * (((tj + 1) - 1.0) * p.ny + ((j + 1) - 0.5));
        flux_sources(d, p, s.tstage, d.ctrl->t, x, y, Qt, psiQt);
        Ew = stw + (0.0 + Qt / (gam * gam));
        Eh = sth + (0.0 + psiQt / gam);
    } else {
        Ew = stw + 0.0;
        Eh = sth + 0.0;
    }
    double hpg = -0.0;
    hpg = hpg + (ow - ob0);
    hpg = hpg / gam;
    const double Eu = stu + (0.0 - p.g * orho * hpg * bx);
    const double Ev = stv + (0.0 - p.g * orho * hpg * by);
    // DragClosure + ImplicitSourceTerms (src/Equations.f90:627-658)
    double Im = 0.0;
    if (ohn > p.hthr) {
        const double fric = drag(p, ohn, ou, ov, opsi, bx, by, gam);
        const double modu = sqrt(speed2(p.geometric, ou, ov, bx, by));
        if (modu > 1.0e-8) {
            const double hr = 1.0 / ohn;
            Im = -(fric * hr / modu);
        }
    }

    if (STAGE == 1) {
        d.R[oidx] = Ew; d.R[FS + oidx] = Eu; d.R[2 * FS + oidx] = Ev; d.R[3 * FS + oidx] = Eh;
        d.R[4 * FS + oidx] = Im;
    } else if (STAGE == 2) {
        // second SSP substep (src/TimeStepper.f90:407-444)
        const double u2 = 0.75 * pru + 0.25 * (oru + dt * Eu) / (1.0 - dt * Im);
        const double v2 = 0.75 * prv + 0.25 * (orv + dt * Ev) / (1.0 - dt * Im);
        double hpo = -0.0;
        hpo = hpo + (pw - ob0);
        double hpn = -0.0;
        hpn = hpn + (ow - ob0);
        double wu = 0.0;
        wu = wu + 0.25 * hpn;
        wu = wu + 0.75 * hpo;
        wu = wu + 0.25 * dt * Ew;
        wu = wu + ob0;
        const double h2 = 0.75 * ph + 0.25 * (oh + dt * Eh);
        d.S2[oidx] = wu; d.S2[FS + oidx] = u2; d.S2[2 * FS + oidx] = v2; d.S2[3 * FS + oidx] = h2;
    } else if (STAGE == 3) {
        // third SSP substep (src/TimeStepper.f90:466-498)
        const double w0 = d.S0[oidx], u0 = d.S0[FS + oidx], v0 = d.S0[2 * FS + oidx];
        const double h0 = d.S0[3 * FS + oidx];
        const double third = 1.0 / 3.0, twothird = 2.0 / 3.0;
        const double u3 = third * u0 + twothird * (oru + dt * Eu) / (1.0 - dt * Im);
        const double v3 = third * v0 + twothird * (orv + dt * Ev) / (1.0 - dt * Im);
        double hpo = -0.0;
        hpo = hpo + (w0 - ob0);
        double hpn = -0.0;
        hpn = hpn + (ow - ob0);
        double wu = 0.0;
        wu = wu + third * hpo;
        wu = wu + twothird * hpn;
        wu = wu + twothird * dt * Ew;
        wu = wu + ob0;
        const double h3 = third * h0 + twothird * (oh + dt * Eh);
        d.S3[oidx] = wu; d.S3[FS + oidx] = u3; d.S3[2 * FS + oidx] = v3; d.S3[3 * FS + oidx] = h3;
    } else {
        // final implicit substep (src/TimeStepper.f90:504-518)
        const double un = (oru - dt * dt * Eu * Im) / (1.0 + dt * dt * Im * Im);
        const double vn = (orv - dt * dt * Ev * Im) / (1.0 + dt * dt * Im * Im);
        if (p.track_maxima) {
            // UpdateMaximum* on the pre-step state at nextT (src/TimeStepper.f90:1155-1303)
            const double w0 = d.S0[oidx], u0 = d.S0[FS + oidx], v0 = d.S0[2 * FS + oidx];
            const double h0 = d.S0[3 * FS + oidx];
            double hn = compute_hn(w0, ob0, 0.0, gam);
            if (hn < 0.0) hn = 0.0;
            if (hn > p.hthr) {
                double hp0 = h0;
                if (hp0 < 0.0) hp0 = 0.0;
                const double eps = p.hthr;
                const double den = hn * hn + dmax(hn * hn, eps * eps);
                const double psi = dmin(2.0 * hn * hp0 / den, p.max_pack);
                const double rho = density(p, psi);
                const double uu = 2.0 * hn * u0 / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
                const double vv = 2.0 * hn * v0 / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
                const double tn = d.ctrl->t + dt;
                double* M = d.maxima;
                if (M[10 * FS + oidx] == -1) M[10 * FS + oidx] = tn;
                if (hn > M[oidx]) { M[oidx] = hn; M[FS + oidx] = tn; }
                const double spd = sqrt(speed2(p.geometric, uu, vv, bx, by));
                if (spd > M[2 * FS + oidx]) { M[2 * FS + oidx] = spd; M[3 * FS + oidx] = tn; }
                if (psi > M[8 * FS + oidx]) { M[8 * FS + oidx] = psi; M[9 * FS + oidx] = tn; }
            }
        }
        d.S0[oidx] = ow; d.S0[FS + oidx] = un; d.S0[2 * FS + oidx] = vn; d.S0[3 * FS + oidx] = oh;
        // activation flags from this state's Hn (CheckIfNearBoundaries,
        // src/TimeStepper.f90:946-1150): bit 0 N, 1 S, 2 E, 3 W
        if (ohn > p.hthr) {
            const int i = s.bx0 + tx, j = s.by0 + ty, tb = p.tile_buffer;
            unsigned f = 0;
            if (j + 1 > p.ny - tb) f |= 1u;
            if (j + 1 <= tb) f |= 2u;
            if (i + 1 > p.nx - tb) f |= 4u;
            if (i + 1 <= tb) f |= 8u;
            if (f) atomicOr(&d.eflags[t], f);
        }
    }
}

// ---------------------------------------------------------------------------
// tile activation (src/TimeStepper.f90:924-1150, src/UpdateTiles.f90:56-481)
// ---------------------------------------------------------------------------

// Ghost w slopes at the edges of tile g that face active tiles
// (ReconstructwAtEdges, src/UpdateTiles.f90:669-750).
__device__ void ghost_slopes(const Dev& d, const Params& p, int g) {
    const int nx = p.nx, ny = p.ny, NC = p.NC;
    const double* w = d.S0;
    const size_t base = (size_t)g * NC;
    auto on = [&](int k) { return k >= 0 && k < p.ntiles && d.tstate[k] == KST_TILE_ACTIVE; };
    int nb = g - 1;
    if (on(nb))
        for (int j = threadIdx.x; j < ny; j += blockDim.x) {
            const double c = w[base + j * nx], e = w[base + j * nx + 1];
            const double wv = w[(size_t)nb * NC + j * nx + nx - 1];
            d.gsx[base + j * nx] = p.dxr * limiter(p.limiter, e - c, c - wv);
        }
    nb = g + 1;
    if (on(nb))
        for (int j = threadIdx.x; j < ny; j += blockDim.x) {
            const double c = w[base + j * nx + nx - 1], we = w[base + j * nx + nx - 2];
            const double ev = w[(size_t)nb * NC + j * nx];
            d.gsx[base + j * nx + nx - 1] = p.dxr * limiter(p.limiter, ev - c, c - we);
        }
    nb = g - p.ntx;
    if (on(nb))
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const double c = w[base + i], n = w[base + nx + i];
            const double sv = w[(size_t)nb * NC + (ny - 1) * nx + i];
            d.gsy[base + i] = p.dyr * limiter(p.limiter, n - c, c - sv);
        }
    nb = g + p.ntx;
    if (on(nb))
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const double c = w[base + (ny - 1) * nx + i], sv = w[base + (ny - 2) * nx + i];
            const double nv = w[(size_t)nb * NC + i];
            d.gsy[base + (ny - 1) * nx + i] = p.dyr * limiter(p.limiter, nv - c, c - sv);
        }
}

// Dry data of a freshly activated or ghost tile (ActivateTile sets w = b0,
// SetDefaultTileData zeroes the rest; src/UpdateTiles.f90:342-384,647-664).
__device__ void dry_tile(const Dev& d, const Params& p, int k, bool active) {
    const size_t base = (size_t)k * p.NC;
    const long long FS = d.FS;
    for (int c = threadIdx.x; c < p.NC; c += blockDim.x) {
        const double b = d.b0c[base + c];
        double* S[3] = {d.S0, d.S2, d.S3};
        for (int q = 0; q < 3; ++q) {
            S[q][base + c] = b;
            S[q][FS + base + c] = 0.0;
            S[q][2 * FS + base + c] = 0.0;
            S[q][3 * FS + base + c] = 0.0;
        }
        for (int q = 0; q < 5; ++q) d.R[q * FS + base + c] = 0.0;
        if (active) {
            for (int q = 0; q < 10; ++q) d.maxima[q * FS + base + c] = 0.0;
            d.maxima[10 * FS + base + c] = -1.0;
        }
    }
}

static constexpr int MAXNEW = 1024;

__global__ __launch_bounds__(256) void activate_kernel(Dev d, Params p) {
    __shared__ int newt[MAXNEW];
    __shared__ int nnew, bad, need;
    Ctrl* c = d.ctrl;
    if (c->status != ST_NEW) return;
    // fast path: does any flagged edge face a tile that is not active?
    if (threadIdx.x == 0) need = 0;
    __syncthreads();
    {
        const int dirs[4] = {p.ntx, -p.ntx, 1, -1};
        int mine = 0;
        for (int q = threadIdx.x; q < c->n_active; q += blockDim.x) {
            const int k = d.active[q];
            const unsigned f = d.eflags[k];
            if (!f) continue;
            for (int r = 0; r < 4; ++r) {
                const int n = k + dirs[r];
                if ((f & (1u << r)) && n > -1 && !(n < p.ntiles && d.tstate[n] == KST_TILE_ACTIVE)) mine = 1;
            }
        }
        if (mine) atomicOr(&need, 1);
    }
    __syncthreads();
    if (!need) {
        for (int q = threadIdx.x; q < c->n_active; q += blockDim.x) d.eflags[d.active[q]] = 0;
        return;
    }
    if (threadIdx.x == 0) {
        nnew = 0;
        bad = 0;
        const int dirs[4] = {p.ntx, -p.ntx, 1, -1};  // N, S, E, W passes in order
        for (int q = 0; q < 4 && !bad; ++q) {
            const int size = c->n_active;
            for (int tt = 0; tt < size && !bad; ++tt) {
                const int k = d.active[tt];
                const int n = k + dirs[q];
                if (n > -1 && !(n < p.ntiles && d.tstate[n] == KST_TILE_ACTIVE) &&
                    (d.eflags[k] & (1u << q))) {
                    // AddTile (src/UpdateTiles.f90:56-78): halt at the domain edge
                    const int gi = n % p.ntx, gj = n / p.ntx;
                    const bool edge = n >= p.ntiles || gi == 0 || gi == p.ntx - 1 ||
                                      (p.nty > 1 && (gj == 0 || gj == p.nty - 1));
                    if (edge) { bad = 1; break; }
                    // sorted insertion; later indices of this pass shift as in the reference
                    int pos = c->n_active;
                    while (pos > 0 && d.active[pos - 1] > n) { d.active[pos] = d.active[pos - 1]; --pos; }
                    d.active[pos] = n;
                    c->n_active++;
                    if (d.tstate[n] == KST_TILE_GHOST) {
                        for (int gq = 0; gq < c->n_ghost; ++gq)
                            if (d.ghosts[gq] == n) {
                                for (int r = gq; r < c->n_ghost - 1; ++r) d.ghosts[r] = d.ghosts[r + 1];
                                c->n_ghost--;
                                break;
                            }
                    }
                    d.tstate[n] = KST_TILE_ACTIVE;
                    d.eflags[n] = 0;
                    if (nnew < MAXNEW) newt[nnew++] = n;
                    else bad = 2;
                }
            }
        }
        if (bad) {
            c->status = ST_ERROR;
            c->err = (bad == 1) ? KST_ERR_HALT : KST_ERR_STATE;
        }
    }
    __syncthreads();
    if (bad) return;
    // clear the flags consumed by this pass
    for (int q = threadIdx.x; q < c->n_active; q += blockDim.x) d.eflags[d.active[q]] = 0;
    if (nnew == 0) return;
    __syncthreads();
    for (int q = 0; q < nnew; ++q) dry_tile(d, p, newt[q], true);
    // new ghosts around each new tile (AddGhostTiles, src/UpdateTiles.f90:389-481)
    __shared__ int ng_start;
    if (threadIdx.x == 0) {
        ng_start = c->n_ghost;
        for (int q = 0; q < nnew; ++q) {
            const int k = newt[q];
            const int nb[8] = {k - 1, k + 1, k + p.ntx, k - p.ntx,
                               k - p.ntx - 1, k - p.ntx + 1, k + p.ntx - 1, k + p.ntx + 1};
            for (int r = 0; r < 8; ++r) {
                const int g = nb[r];
                if (g >= 0 && g < p.ntiles && d.tstate[g] == KST_TILE_OFF) {
                    d.tstate[g] = KST_TILE_GHOST;
                    d.ghosts[c->n_ghost++] = g;
                }
            }
        }
    }
    __syncthreads();
    for (int q = ng_start; q < c->n_ghost; ++q) dry_tile(d, p, d.ghosts[q], false);
    __syncthreads();
    // ghosts whose neighbourhood changed rebuild their edge w slopes
    for (int q = 0; q < c->n_ghost; ++q) {
        const int g = d.ghosts[q];
        bool touched = q >= ng_start;
        for (int r = 0; r < nnew && !touched; ++r) {
            const int k = newt[r];
            const int di = (g % p.ntx) - (k % p.ntx), dj = (g / p.ntx) - (k / p.ntx);
            if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1) touched = true;
        }
        if (touched) ghost_slopes(d, p, g);
    }
}

// Rebuild every ghost's edge slopes (end of LoadSourceConditions,
// src/SetSources.f90:388-392) and the activation flags of every active tile.
__global__ __launch_bounds__(256) void init_ghosts_kernel(Dev d, Params p) {
    const int q = blockIdx.x;
    if (q >= d.ctrl->n_ghost) return;
    ghost_slopes(d, p, d.ghosts[q]);
}

__global__ __launch_bounds__(256) void init_flags_kernel(Dev d, Params p) {
    const int q = blockIdx.x;
    if (q >= d.ctrl->n_active) return;
    const int t = d.active[q];
    unsigned f = 0;
    const int tb = p.tile_buffer;
    for (int c = threadIdx.x; c < p.NC; c += blockDim.x) {
        const size_t gi = (size_t)t * p.NC + c;
        double hn = compute_hn(d.S0[gi], d.b0c[gi], 0.0, d.cgam[gi]);
        if (hn > p.hthr) {
            const int i = c % p.nx, j = c / p.nx;
            if (j + 1 > p.ny - tb) f |= 1u;
            if (j + 1 <= tb) f |= 2u;
            if (i + 1 > p.nx - tb) f |= 4u;
            if (i + 1 <= tb) f |= 8u;
        }
    }
    if (f) atomicOr(&d.eflags[t], f);
}

// Accept or roll back the step (src/TimeStepper.f90:173-275).
__global__ void finalize_kernel(Dev d, Params p) {
    Ctrl* c = d.ctrl;
    if (c->status == ST_DONE || c->status == ST_ERROR) return;
    const bool ran_p1 = (c->status == ST_NEW);
    const double dt = step_dt(d, p);
    c->rhs_evals += (ran_p1 ? 1 : 0) + 1;
    const double dt1 = advised_from_slot(p, c->slot[1], c->n_active);
    bool reject = false;
    if (dt1 < dt) {
        reject = true;
        c->dt_retry = 0.9 * dt1;
    } else {
        c->rhs_evals += 1;
        const double dt2 = advised_from_slot(p, c->slot[2], c->n_active);
        if (dt2 < dt) {
            reject = true;
            c->dt_retry = 0.9 * dt2;
        } else {
            c->rhs_evals += 1;
        }
    }
    const unsigned long long huge = dbits(kHuge);
    if (reject) {
        c->rejected++;
        c->status = ST_RETRY;
        c->slot[1] = huge;
        c->slot[2] = huge;
        return;
    }
    c->t = c->t + dt;
    c->steps++;
    c->cell_updates += (long long)c->n_active * p.NC;
    c->last_dt = dt;
    if (c->steps == 1 || dt < c->min_dt) c->min_dt = dt;
    if (dt > c->max_dt) c->max_dt = dt;
    c->slot[0] = huge; c->slot[1] = huge; c->slot[2] = huge;
    c->status = ST_NEW;
    if (c->mode == 0) {
        if (c->t >= c->tend) c->status = ST_DONE;
    } else {
        if (--c->steps_left <= 0) c->status = ST_DONE;
    }
}

// Cell and face topography from the tile's vertex elevations
// (ComputeCellCentredTopographicData / ComputeInterfacialTopographicData,
// src/MorphodynamicRHS.f90:308-578; face b0 of GetHeights, src/dem.f90:378-386).
__global__ __launch_bounds__(256) void topo_kernel(Dev d, Params p) {
    const int t = blockIdx.x;
    const int nx = p.nx, ny = p.ny, NV1 = nx + 1;
    const int ti = t % p.ntx, tj = t / p.ntx;
    // neighbour tiles (clamped at the domain edge: their values are never read)
    const int tW = ti > 0 ? t - 1 : t, tE = ti < p.ntx - 1 ? t + 1 : t;
    const int tS = tj > 0 ? t - p.ntx : t, tN = tj < p.nty - 1 ? t + p.ntx : t;
    const size_t NV = (size_t)(nx + 1) * (ny + 1);
    const double* b = d.vb0 + (size_t)t * NV;
    const double* bt = d.vbt + (size_t)t * NV;
    const double* bW = d.vb0 + (size_t)tW * NV; const double* btW = d.vbt + (size_t)tW * NV;
    const double* bE = d.vb0 + (size_t)tE * NV; const double* btE = d.vbt + (size_t)tE * NV;
    const double* bS = d.vb0 + (size_t)tS * NV; const double* btS = d.vbt + (size_t)tS * NV;
    const double* bN = d.vb0 + (size_t)tN * NV; const double* btN = d.vbt + (size_t)tN * NV;
#define V(a, i, j) a[(j) * NV1 + (i)]
    for (int c = threadIdx.x; c < p.NC; c += blockDim.x) {
        const int i = c % nx, j = c / nx;
        Kahan k;
        k.add(V(b, i, j)); k.add(V(b, i + 1, j)); k.add(V(b, i, j + 1)); k.add(V(b, i + 1, j + 1));
        const double b0c = 0.25 * k.s;
        Kahan kt;
        kt.add(V(bt, i, j)); kt.add(V(bt, i + 1, j)); kt.add(V(bt, i, j + 1)); kt.add(V(bt, i + 1, j + 1));
        const double btc = 0.25 * kt.s;
        Kahan kx;
        kx.add(V(b, i + 1, j)); kx.add(V(bt, i + 1, j)); kx.add(-V(b, i, j)); kx.add(-V(bt, i, j));
        kx.add(V(b, i + 1, j + 1)); kx.add(V(bt, i + 1, j + 1)); kx.add(-V(b, i, j + 1)); kx.add(-V(bt, i, j + 1));
        const double dbdx = 0.5 * p.dxr * kx.s;
        Kahan ky;
        ky.add(V(b, i, j + 1)); ky.add(V(bt, i, j + 1)); ky.add(-V(b, i, j)); ky.add(-V(bt, i, j));
        ky.add(V(b, i + 1, j + 1)); ky.add(V(bt, i + 1, j + 1)); ky.add(-V(b, i + 1, j)); ky.add(-V(bt, i + 1, j));
        const double dbdy = 0.5 * p.dyr * ky.s;
        const size_t gi = (size_t)t * p.NC + c;
        d.b0c[gi] = b0c; d.btc[gi] = btc; d.cbx[gi] = dbdx; d.cby[gi] = dbdy;
        d.cgam[gi] = gamma_of(p.geometric, dbdx, dbdy);
    }
    // x-faces: i in [0, nx], j in [0, ny)
    for (int f = threadIdx.x; f < (nx + 1) * ny; f += blockDim.x) {
        const int i = f % (nx + 1), j = f / (nx + 1);
        Kahan kx;
        if (i >= 1 && i <= nx - 1) {
            kx.add(V(b, i + 1, j)); kx.add(V(bt, i + 1, j)); kx.add(V(b, i + 1, j + 1)); kx.add(V(bt, i + 1, j + 1));
            kx.add(-V(b, i - 1, j)); kx.add(-V(bt, i - 1, j)); kx.add(-V(b, i - 1, j + 1)); kx.add(-V(bt, i - 1, j + 1));
        } else if (i == 0) {
            kx.add(V(b, 1, j)); kx.add(V(bt, 1, j)); kx.add(V(b, 1, j + 1)); kx.add(V(bt, 1, j + 1));
            kx.add(-V(bW, nx - 1, j)); kx.add(-V(btW, nx - 1, j)); kx.add(-V(bW, nx - 1, j + 1)); kx.add(-V(btW, nx - 1, j + 1));
        } else {
            kx.add(V(bE, 1, j)); kx.add(V(btE, 1, j)); kx.add(V(bE, 1, j + 1)); kx.add(V(btE, 1, j + 1));
            kx.add(-V(b, nx - 1, j)); kx.add(-V(bt, nx - 1, j)); kx.add(-V(b, nx - 1, j + 1)); kx.add(-V(bt, nx - 1, j + 1));
        }
        const double dbdx = 0.25 * p.dxr * kx.s;
        const double btf = 0.5 * (V(bt, i, j) + V(bt, i, j + 1));
        Kahan ky;
        ky.add(V(b, i, j + 1)); ky.add(V(bt, i, j + 1)); ky.add(-V(b, i, j)); ky.add(-V(bt, i, j));
        const double dbdy = p.dyr * ky.s;
        const size_t fi = (size_t)t * p.NFX + f;
        d.xb0[fi] = 0.5 * (V(b, i, j) + V(b, i, j + 1));
        d.xbt[fi] = btf;
        d.xby[fi] = dbdy;
        d.xgam[fi] = gamma_of(p.geometric, dbdx, dbdy);
    }
    // y-faces: i in [0, nx), j in [0, ny]
    for (int f = threadIdx.x; f < nx * (ny + 1); f += blockDim.x) {
        const int i = f % nx, j = f / nx;
        Kahan kx;
        kx.add(V(b, i + 1, j)); kx.add(V(bt, i + 1, j)); kx.add(-V(b, i, j)); kx.add(-V(bt, i, j));
        const double dbdx = p.dxr * kx.s;
        const double btf = 0.5 * (V(bt, i, j) + V(bt, i + 1, j));
        Kahan ky;
        if (j >= 1 && j <= ny - 1) {
            ky.add(V(b, i + 1, j + 1)); ky.add(V(bt, i + 1, j + 1)); ky.add(V(b, i, j + 1)); ky.add(V(bt, i, j + 1));
            ky.add(-V(b, i + 1, j - 1)); ky.add(-V(bt, i + 1, j - 1)); ky.add(-V(b, i, j - 1)); ky.add(-V(bt, i, j - 1));
        } else if (j == 0) {
            ky.add(V(b, i + 1, 1)); ky.add(V(bt, i + 1, 1)); ky.add(V(b, i, 1)); ky.add(V(bt, i, 1));
            ky.add(-V(bS, i + 1, ny - 1)); ky.add(-V(btS, i + 1, ny - 1)); ky.add(-V(bS, i, ny - 1)); ky.add(-V(btS, i, ny - 1));
        } else {
            ky.add(V(bN, i + 1, 1)); ky.add(V(btN, i + 1, 1)); ky.add(V(bN, i, 1)); ky.add(V(btN, i, 1));
            ky.add(-V(b, i + 1, ny - 1)); ky.add(-V(bt, i + 1, ny - 1)); ky.add(-V(b, i, ny - 1)); ky.add(-V(bt, i, ny - 1));
        }
        const double dbdy = 0.25 * p.dyr * ky.s;
        const size_t fi = (size_t)t * p.NFY + f;
        d.yb0[fi] = 0.5 * (V(b, i, j) + V(b, i + 1, j));
        d.ybt[fi] = btf;
        d.ybx[fi] = dbdx;
        d.ygam[fi] = gamma_of(p.geometric, dbdx, dbdy);
    }
#undef V
}

// Static per-face CFL numerators and per-cell surface-gradient prefactors.
// gam_ratio = min(gam(cell)/gam(face), 1) of the cell on each side of a face
// (src/HydraulicRHS.f90:983-1006) enters the CFL as (gr*gr)*dx; the
// prefactors of ConstructHydraulicRHS (src/HydraulicRHS.f90:1233-1243).
// Needs cgam of the neighbour tiles, so it runs after topo_kernel.
__global__ __launch_bounds__(256) void topo2_kernel(Dev d, Params p) {
    const int t = blockIdx.x;
    const int nx = p.nx, ny = p.ny;
    const int ti = t % p.ntx, tj = t / p.ntx;
    const int tW = ti > 0 ? t - 1 : t, tE = ti < p.ntx - 1 ? t + 1 : t;
    const int tS = tj > 0 ? t - p.ntx : t, tN = tj < p.nty - 1 ? t + p.ntx : t;
    for (int c = threadIdx.x; c < p.NC; c += blockDim.x) {
        const int gi = t * p.NC + c;
        const double bx = d.cbx[gi], by = d.cby[gi], gam = d.cgam[gi];
        if (p.geometric) {
            d.pxu[gi] = (1.0 + by * by) / gam;
            d.pxv[gi] = -(bx * by / gam);
            d.pyv[gi] = (1.0 + bx * bx) / gam;
        } else {
            d.pxu[gi] = 1.0; d.pxv[gi] = 0.0; d.pyv[gi] = 1.0;
        }
    }
    for (int f = threadIdx.x; f < (nx + 1) * ny; f += blockDim.x) {
        const int i = f % (nx + 1), j = f / (nx + 1);
        const double fg = d.xgam[t * p.NFX + f];
        const double gl = (i == 0) ? d.cgam[tW * p.NC + j * nx + nx - 1] : d.cgam[t * p.NC + j * nx + i - 1];
        const double gr_ = (i == nx) ? d.cgam[tE * p.NC + j * nx] : d.cgam[t * p.NC + j * nx + i];
        const double a = fmin(gl / fg, 1.0), b = fmin(gr_ / fg, 1.0);
        d.xnp[t * p.NFX + f] = a * a * p.dx;
        d.xnm[t * p.NFX + f] = b * b * p.dx;
    }
    for (int f = threadIdx.x; f < nx * (ny + 1); f += blockDim.x) {
        const int i = f % nx, j = f / nx;
        const double fg = d.ygam[t * p.NFY + f];
        const double gl = (j == 0) ? d.cgam[tS * p.NC + (ny - 1) * nx + i] : d.cgam[t * p.NC + (j - 1) * nx + i];
        const double gr_ = (j == ny) ? d.cgam[tN * p.NC + i] : d.cgam[t * p.NC + j * nx + i];
        const double a = fmin(gl / fg, 1.0), b = fmin(gr_ / fg, 1.0);
        d.ynp[t * p.NFY + f] = a * a * p.dy;
        d.ynm[t * p.NFY + f] = b * b * p.dy;
    }
}

}  // namespace kst

// ===========================================================================
// host side: the C-ABI (include/kestrel_step.h)
// ===========================================================================
using namespace kst;

struct kst_handle {
    kst_config cfg;
    Params p;
    Dev d;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    std::string err;
    int nx, ny, ntx, nty, ntiles, NX, NY, NC;
    bool have_topo = false, have_state = false;
    double t = 0.0;
    // profiling: per-iteration event pairs around the four stage kernels
    bool profiling = false;
    hipGraphExec_t gexec = nullptr;      // graph of kGraphIters whole iterations
    std::vector<hipEvent_t> ev;          // 8 per iteration of the current batch
    double prof_ms[4] = {0, 0, 0, 0};
    long long prof_n[4] = {0, 0, 0, 0};
};

namespace {

int set_err(kst_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

#define HIPCHK(h, call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return set_err(h, KST_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(kst_handle* h, T** ptr, size_t n) {
    void* p = nullptr;
    HIPCHK(h, hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    HIPCHK(h, hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), h->stream));
    h->allocs.push_back(p);
    *ptr = static_cast<T*>(p);
    return 0;
}

void fill_params(kst_handle* h) {
    const kst_config& c = h->cfg;
    Params& p = h->p;
    std::memset(&p, 0, sizeof p);
    p.nx = c.nx_per_tile; p.ny = c.ny_per_tile; p.ntx = c.nx_tiles; p.nty = c.ny_tiles;
    p.ntiles = c.nx_tiles * c.ny_tiles;
    p.NC = p.nx * p.ny; p.NFX = (p.nx + 1) * p.ny; p.NFY = p.nx * (p.ny + 1);
    p.limiter = c.limiter; p.geometric = c.geometric_factors; p.drag = c.drag;
    p.fswitch = c.fswitch; p.tile_buffer = c.tile_buffer; p.track_maxima = c.track_maxima;
    p.n_sources = c.n_sources; p.erosion = c.erosion; p.deposition = c.deposition;
    p.ero_transition = c.ero_transition; p.morpho_damping = c.morpho_damping;
    p.dx = c.dx; p.dy = c.dy; p.dxr = 1.0 / c.dx; p.dyr = 1.0 / c.dy;
    p.x_size = c.x_size; p.y_size = c.y_size;
    p.g = c.g; p.rhow = c.rhow; p.rhos = c.rhos;
    p.chezy = c.chezy_co; p.coulomb = c.coulomb_co; p.manning = c.manning_co;
    p.pmin = c.pouliquen_min; p.pmax = c.pouliquen_max; p.pint = c.pouliquen_intermediate;
    p.pbeta = c.pouliquen_beta; p.ebst = c.edwards_betastar; p.ekap = c.edwards_kappa;
    p.egam = c.edwards_gamma; p.ero_rate = c.ero_rate; p.ero_rate_granular = c.ero_rate_granular;
    p.ero_depth = c.ero_depth; p.ero_hcrit = c.ero_critical_height;
    p.vrate = c.voellmy_switch_rate; p.vvalue = c.voellmy_switch_value;
    p.max_pack = c.max_pack; p.porosity = c.bed_porosity; p.diam = c.solid_diameter;
    p.nu = c.eddy_viscosity; p.cfl = c.cfl; p.max_dt = c.max_dt; p.hthr = c.height_threshold;
    p.diff_ts = c.diffusive_time_scale; p.gred = c.gred; p.ws0 = c.ws0;
    p.nsettling = c.nsettling; p.crit_shields = c.critical_shields;
}

int read_ctrl(kst_handle* h, Ctrl* c) {
    HIPCHK(h, hipMemcpyAsync(c, h->d.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

int write_ctrl(kst_handle* h, const Ctrl* c) {
    HIPCHK(h, hipMemcpyAsync(h->d.ctrl, c, sizeof(Ctrl), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
}

// Global row-major (NY, NX) <-> tile-major [tile][ny][nx].
void to_tiles(const kst_handle* h, const double* g, double* t) {
    for (int J = 0; J < h->NY; ++J)
        for (int I = 0; I < h->NX; ++I) {
            const int ti = I / h->nx, tj = J / h->ny, i = I % h->nx, j = J % h->ny;
            t[(size_t)(ti + tj * h->ntx) * h->NC + j * h->nx + i] = g[(size_t)J * h->NX + I];
        }
}
void from_tiles(const kst_handle* h, const double* t, double* g) {
    for (int J = 0; J < h->NY; ++J)
        for (int I = 0; I < h->NX; ++I) {
            const int ti = I / h->nx, tj = J / h->ny, i = I % h->nx, j = J % h->ny;
            g[(size_t)J * h->NX + I] = t[(size_t)(ti + tj * h->ntx) * h->NC + j * h->nx + i];
        }
}

typedef void (*StageFn)(Dev, Params);
struct StageSet { StageFn fn[4]; };

template <bool VISC, int LIM>
StageSet make_set() {
    return StageSet{{stage_kernel<1, VISC, LIM>, stage_kernel<2, VISC, LIM>,
                     stage_kernel<3, VISC, LIM>, stage_kernel<4, VISC, LIM>}};
}

const StageSet& stage_set(bool visc, int limiter) {
    static const StageSet sets[2][5] = {
        {make_set<false, 0>(), make_set<false, 1>(), make_set<false, 2>(), make_set<false, 3>(), make_set<false, 4>()},
        {make_set<true, 0>(), make_set<true, 1>(), make_set<true, 2>(), make_set<true, 3>(), make_set<true, 4>()}};
    return sets[visc ? 1 : 0][(limiter >= 0 && limiter < 5) ? limiter : 1];
}

int launch_iteration(kst_handle* h, int slot) {
    const Params& p = h->p;
    const Dev& d = h->d;
    hipEvent_t* E = (h->profiling && slot >= 0) ? &h->ev[8 * slot] : nullptr;
#define EVREC(k) if (E) HIPCHK(h, hipEventRecord(E[k], h->stream))
    const int spt = ((p.nx + B - 1) / B) * ((p.ny + B - 1) / B);
    // 8 * ceil(work/8) blocks so the XCD-chunked mapping is a bijection
    const int nblk = 8 * ((p.ntiles * spt + 7) / 8);
    const bool visc = p.nu != 0.0;
    hipLaunchKernelGGL(activate_kernel, dim3(1), dim3(256), 0, h->stream, d, p);
    const StageSet& k = stage_set(visc, p.limiter);
    for (int q = 0; q < 4; ++q) {
        EVREC(2 * q);
        hipLaunchKernelGGL(k.fn[q], dim3(nblk), dim3(NTHREADS), 0, h->stream, d, p);
        EVREC(2 * q + 1);
    }
#undef EVREC
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(1), 0, h->stream, d, p);
    HIPCHK(h, hipGetLastError());
    return 0;
}

constexpr int kGraphIters = 8;

// One hipGraph holding kGraphIters whole iterations (activation, four stage
// kernels, finalize); kernel arguments never change for a handle.
int ensure_graph(kst_handle* h) {
    if (h->gexec) return 0;
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int q = 0; q < kGraphIters; ++q) {
        int rc = launch_iteration(h, -1);
        if (rc) {
            hipGraph_t dummy = nullptr;
            (void)hipStreamEndCapture(h->stream, &dummy);
            if (dummy) (void)hipGraphDestroy(dummy);
            return rc;
        }
    }
    HIPCHK(h, hipStreamEndCapture(h->stream, &g));
    HIPCHK(h, hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0));
    HIPCHK(h, hipGraphDestroy(g));
    return 0;
}

int run_until_done(kst_handle* h, kst_stats* st, long long max_iters) {
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    int batch = h->profiling ? 4 : kGraphIters;
    long long launched = 0;
    for (;;) {
        if (launched > max_iters)
            return set_err(h, KST_ERR_TOO_MANY_STEPS, "step limit exceeded before reaching the target");
        launched += batch;
        const long long done_before = c.steps + c.rejected;
        const long long new_before = c.steps;
        if (h->profiling && (int)h->ev.size() < 8 * batch) {
            const size_t old = h->ev.size();
            h->ev.resize(8 * batch);
            for (size_t q = old; q < h->ev.size(); ++q) HIPCHK(h, hipEventCreate(&h->ev[q]));
        }
        if (h->profiling) {
            for (int q = 0; q < batch; ++q) {
                rc = launch_iteration(h, q);
                if (rc) return rc;
            }
        } else {
            rc = ensure_graph(h);
            if (rc) return rc;
            for (int q = 0; q < batch; q += kGraphIters) HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
        }
        rc = read_ctrl(h, &c);
        if (rc) return rc;
        if (h->profiling) {
            // the first (steps + rejected) iterations of the batch did work;
            // P1 ran only in iterations that started a new step
            const long long productive = c.steps + c.rejected - done_before;
            long long p1 = c.steps - new_before;
            for (long long q = 0; q < productive && q < batch; ++q)
                for (int k = 0; k < 4; ++k) {
                    if (k == 0 && p1 <= 0) continue;
                    float ms = 0.f;
                    HIPCHK(h, hipEventElapsedTime(&ms, h->ev[8 * q + 2 * k], h->ev[8 * q + 2 * k + 1]));
                    h->prof_ms[k] += ms;
                    h->prof_n[k] += 1;
                    if (k == 0) --p1;
                }
        }
        if (c.status == ST_ERROR) {
            if (c.err == KST_ERR_HALT)
                return set_err(h, KST_ERR_HALT,
                               "Error: tried to add a tile outside the domain. Try increasing the "
                               "domain size or repositioning its location.");
            return set_err(h, c.err ? c.err : KST_ERR_STATE, "device error");
        }
        if (c.status == ST_DONE) break;
        batch = std::min(batch * 2, 64);
    }
    h->t = c.t;
    if (st) {
        st->steps = c.steps; st->rejected = c.rejected; st->rhs_evals = c.rhs_evals;
        st->cell_updates = c.cell_updates; st->active_tiles = c.n_active; st->ghost_tiles = c.n_ghost;
        st->t = c.t; st->last_dt = c.last_dt; st->min_dt = c.min_dt; st->max_dt = c.max_dt;
    }
    return 0;
}

}  // namespace

extern "C" {

void kst_config_defaults(kst_config* c) {
    std::memset(c, 0, sizeof *c);
    c->bc = KST_BC_HALT;
    c->limiter = KST_LIMITER_MINMOD2;
    c->geometric_factors = 1;
    c->drag = KST_DRAG_CHEZY;
    c->erosion = KST_EROSION_MIXED;
    c->deposition = KST_DEPOSITION_SPEARMAN_MANNING;
    c->ero_transition = KST_TRANSITION_SMOOTH;
    c->morpho_damping = KST_DAMPING_TANH;
    c->fswitch = KST_SWITCH_TANH;
    c->morphodynamics = 1;
    c->tile_buffer = 1;
    c->track_maxima = 1;
    c->precision = KST_PREC_FP64;
    c->g = 9.81; c->rhow = 1000.0; c->rhos = 2000.0;
    c->chezy_co = 0.01; c->coulomb_co = 0.1; c->manning_co = 0.03;
    c->pouliquen_min = 0.1; c->pouliquen_max = 0.4; c->pouliquen_intermediate = 0.2;
    c->pouliquen_beta = 0.136; c->edwards_betastar = 0.136; c->edwards_kappa = 1.0;
    c->edwards_gamma = 0.0; c->ero_rate = 0.001; c->ero_rate_granular = 4.0; c->ero_depth = 1.0;
    c->ero_critical_height = 0.01; c->voellmy_switch_rate = 3.0; c->voellmy_switch_value = 0.2;
    c->max_pack = 0.65; c->bed_porosity = 0.35; c->solid_diameter = 1e-3; c->eddy_viscosity = 0.0;
    c->cfl = 0.25; c->max_dt = DBL_MAX; c->height_threshold = 1e-6;
    kst_config_derive(c);
}

void kst_config_derive(kst_config* c) {
    const double visc_w = 1.2e-6;
    c->gred = (c->rhos / c->rhow - 1.0) * c->g;
    c->rep = std::sqrt(c->g * c->solid_diameter) * c->solid_diameter / visc_w;
    const double R = std::pow(c->gred / visc_w / visc_w, 1.0 / 3.0) * c->solid_diameter;
    c->ws0 = visc_w / c->solid_diameter * (std::sqrt(10.36 * 10.36 + 1.048 * R * R * R) - 10.36);
    c->nsettling = (4.7 + 0.41 * std::pow(c->rep, 0.75)) / (1.0 + 0.175 * std::pow(c->rep, 0.75));
    c->critical_shields = 0.3 / (1.0 + 1.2 * R) + 0.055 * (1.0 - std::exp(-0.02 * R));
    c->diffusive_time_scale = DBL_MAX;
    if (c->eddy_viscosity > 0.0)
        c->diffusive_time_scale = std::min(c->dx * c->dx / c->eddy_viscosity, c->dy * c->dy / c->eddy_viscosity);
}

int kst_create(const kst_config* cfg, kst_handle** out) {
    if (!cfg || !out) return KST_ERR_CONFIG;
    *out = nullptr;
    if (cfg->nx_per_tile < 4 || cfg->ny_per_tile < 4 || cfg->nx_tiles < 3 || cfg->ny_tiles < 3)
        return KST_ERR_CONFIG;
    if (cfg->bc != KST_BC_HALT || cfg->morphodynamics || cfg->precision != KST_PREC_FP64)
        return KST_ERR_CONFIG;
    kst_handle* h = new kst_handle();
    h->cfg = *cfg;
    h->cfg.sources = nullptr;
    h->nx = cfg->nx_per_tile; h->ny = cfg->ny_per_tile;
    h->ntx = cfg->nx_tiles; h->nty = cfg->ny_tiles;
    h->ntiles = h->ntx * h->nty; h->NC = h->nx * h->ny;
    h->NX = h->ntx * h->nx; h->NY = h->nty * h->ny;
    fill_params(h);
    int rc = 0;
    if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
        delete h;
        return KST_ERR_HIP;
    }
    Dev& d = h->d;
    const size_t FS = (size_t)h->ntiles * h->NC;
    const size_t NV = (size_t)(h->nx + 1) * (h->ny + 1);
    d.FS = (long long)FS;
    rc |= dalloc(h, &d.S0, 4 * FS); rc |= dalloc(h, &d.S2, 4 * FS); rc |= dalloc(h, &d.S3, 4 * FS);
    rc |= dalloc(h, &d.R, 5 * FS);
    rc |= dalloc(h, &d.b0c, FS); rc |= dalloc(h, &d.btc, FS); rc |= dalloc(h, &d.cbx, FS);
    rc |= dalloc(h, &d.cby, FS); rc |= dalloc(h, &d.cgam, FS);
    rc |= dalloc(h, &d.xb0, (size_t)h->ntiles * h->p.NFX); rc |= dalloc(h, &d.xbt, (size_t)h->ntiles * h->p.NFX);
    rc |= dalloc(h, &d.xby, (size_t)h->ntiles * h->p.NFX); rc |= dalloc(h, &d.xgam, (size_t)h->ntiles * h->p.NFX);
    rc |= dalloc(h, &d.yb0, (size_t)h->ntiles * h->p.NFY); rc |= dalloc(h, &d.ybt, (size_t)h->ntiles * h->p.NFY);
    rc |= dalloc(h, &d.ybx, (size_t)h->ntiles * h->p.NFY); rc |= dalloc(h, &d.ygam, (size_t)h->ntiles * h->p.NFY);
    rc |= dalloc(h, &d.vb0, (size_t)h->ntiles * NV); rc |= dalloc(h, &d.vbt, (size_t)h->ntiles * NV);
    rc |= dalloc(h, &d.xnp, (size_t)h->ntiles * h->p.NFX); rc |= dalloc(h, &d.xnm, (size_t)h->ntiles * h->p.NFX);
    rc |= dalloc(h, &d.ynp, (size_t)h->ntiles * h->p.NFY); rc |= dalloc(h, &d.ynm, (size_t)h->ntiles * h->p.NFY);
    rc |= dalloc(h, &d.pxu, FS); rc |= dalloc(h, &d.pxv, FS); rc |= dalloc(h, &d.pyv, FS);
    rc |= dalloc(h, &d.gsx, FS); rc |= dalloc(h, &d.gsy, FS);
    rc |= dalloc(h, &d.maxima, 11 * FS);
    rc |= dalloc(h, &d.tstate, (size_t)h->ntiles); rc |= dalloc(h, &d.has_src, (size_t)h->ntiles);
    rc |= dalloc(h, &d.active, (size_t)h->ntiles + 1); rc |= dalloc(h, &d.ghosts, (size_t)h->ntiles + 1);
    rc |= dalloc(h, &d.eflags, (size_t)h->ntiles); rc |= dalloc(h, &d.ctrl, 1);
    // flux sources
    int ns = cfg->n_sources, nk = 0;
    for (int i = 0; i < ns; ++i) nk += cfg->sources[i].n_series;
    rc |= dalloc(h, &d.src_x, ns); rc |= dalloc(h, &d.src_y, ns); rc |= dalloc(h, &d.src_r, ns);
    rc |= dalloc(h, &d.src_n, ns); rc |= dalloc(h, &d.src_off, ns); rc |= dalloc(h, &d.src_cells, ns);
    rc |= dalloc(h, &d.src_t, nk); rc |= dalloc(h, &d.src_q, nk); rc |= dalloc(h, &d.src_p, nk);
    if (rc) { kst_destroy(h); return KST_ERR_HIP; }
    if (ns > 0) {
        std::vector<double> x(ns), y(ns), r(ns), tt, q, ps;
        std::vector<int> n(ns), off(ns), cells(ns);
        for (int i = 0; i < ns; ++i) {
            const kst_flux_source& s = cfg->sources[i];
            x[i] = s.x; y[i] = s.y; r[i] = s.radius; n[i] = s.n_series; cells[i] = s.num_cells;
            off[i] = (int)tt.size();
            for (int k = 0; k < s.n_series; ++k) { tt.push_back(s.time[k]); q.push_back(s.flux[k]); ps.push_back(s.psi[k]); }
        }
        hipMemcpy(d.src_x, x.data(), 8 * ns, hipMemcpyHostToDevice);
        hipMemcpy(d.src_y, y.data(), 8 * ns, hipMemcpyHostToDevice);
        hipMemcpy(d.src_r, r.data(), 8 * ns, hipMemcpyHostToDevice);
        hipMemcpy(d.src_n, n.data(), 4 * ns, hipMemcpyHostToDevice);
        hipMemcpy(d.src_off, off.data(), 4 * ns, hipMemcpyHostToDevice);
        hipMemcpy(d.src_cells, cells.data(), 4 * ns, hipMemcpyHostToDevice);
        hipMemcpy(d.src_t, tt.data(), 8 * nk, hipMemcpyHostToDevice);
        hipMemcpy(d.src_q, q.data(), 8 * nk, hipMemcpyHostToDevice);
        hipMemcpy(d.src_p, ps.data(), 8 * nk, hipMemcpyHostToDevice);
    }
    *out = h;
    return KST_OK;
}

int kst_set_topography(kst_handle* h, const double* b0v) {
    if (!h || !b0v) return KST_ERR_CONFIG;
    const int nx = h->nx, ny = h->ny, NV1 = nx + 1;
    const size_t NV = (size_t)(nx + 1) * (ny + 1);
    std::vector<double> tv((size_t)h->ntiles * NV);
    for (int tj = 0; tj < h->nty; ++tj)
        for (int ti = 0; ti < h->ntx; ++ti) {
            double* dst = &tv[(size_t)(ti + tj * h->ntx) * NV];
            for (int j = 0; j <= ny; ++j)
                for (int i = 0; i <= nx; ++i)
                    dst[j * NV1 + i] = b0v[(size_t)(tj * ny + j) * (h->NX + 1) + ti * nx + i];
        }
    HIPCHK(h, hipMemcpy(h->d.vb0, tv.data(), tv.size() * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(topo_kernel, dim3(h->ntiles), dim3(256), 0, h->stream, h->d, h->p);
    hipLaunchKernelGGL(topo2_kernel, dim3(h->ntiles), dim3(256), 0, h->stream, h->d, h->p);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->have_topo = true;
    return KST_OK;
}

int kst_set_state(kst_handle* h, const double* w, const double* rhu, const double* rhv,
                  const double* hnpsi, const double* btv, const uint8_t* ts, double t,
                  const double* hnmax0) {
    if (!h || !w || !rhu || !rhv || !hnpsi || !ts) return KST_ERR_CONFIG;
    if (!h->have_topo) return set_err(h, KST_ERR_STATE, "kst_set_topography must come first");
    (void)btv;
    const size_t FS = (size_t)h->ntiles * h->NC;
    // tile lists: active sorted by id; ghosts = non-active 8-neighbours of active tiles
    std::vector<uint8_t> st(h->ntiles, KST_TILE_OFF);
    std::vector<int> act, gh;
    for (int k = 0; k < h->ntiles; ++k)
        if (ts[k] == KST_TILE_ACTIVE) {
            const int gi = k % h->ntx, gj = k / h->ntx;
            if (gi == 0 || gi == h->ntx - 1 || gj == 0 || gj == h->nty - 1)
                return set_err(h, KST_ERR_HALT, "Error: tried to add a tile outside the domain.");
            st[k] = KST_TILE_ACTIVE;
            act.push_back(k);
        }
    for (int k : act) {
        const int gi = k % h->ntx, gj = k / h->ntx;
        for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di) {
                const int n = (gi + di) + (gj + dj) * h->ntx;
                if (st[n] == KST_TILE_OFF) { st[n] = KST_TILE_GHOST; gh.push_back(n); }
            }
    }
    // fields: active tiles take the given state, ghosts take dry data
    std::vector<double> b0c(FS);
    HIPCHK(h, hipMemcpy(b0c.data(), h->d.b0c, FS * 8, hipMemcpyDeviceToHost));
    std::vector<double> S(4 * FS, 0.0);
    to_tiles(h, w, S.data());
    to_tiles(h, rhu, S.data() + FS);
    to_tiles(h, rhv, S.data() + 2 * FS);
    to_tiles(h, hnpsi, S.data() + 3 * FS);
    for (int k = 0; k < h->ntiles; ++k) {
        if (st[k] == KST_TILE_ACTIVE) continue;
        for (int c = 0; c < h->NC; ++c) {
            const size_t gi = (size_t)k * h->NC + c;
            S[gi] = b0c[gi];
            S[FS + gi] = 0.0; S[2 * FS + gi] = 0.0; S[3 * FS + gi] = 0.0;
        }
    }
    HIPCHK(h, hipMemcpy(h->d.S0, S.data(), S.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d.S2, S.data(), S.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d.S3, S.data(), S.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(h->d.R, 0, 5 * FS * 8));
    HIPCHK(h, hipMemset(h->d.gsx, 0, FS * 8));
    HIPCHK(h, hipMemset(h->d.gsy, 0, FS * 8));
    // maxima
    std::vector<double> M(11 * FS, 0.0);
    for (size_t i = 0; i < FS; ++i) M[10 * FS + i] = -1.0;
    if (hnmax0) to_tiles(h, hnmax0, M.data());
    HIPCHK(h, hipMemcpy(h->d.maxima, M.data(), M.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d.tstate, st.data(), st.size(), hipMemcpyHostToDevice));
    std::vector<int> actp(act), ghp(gh);
    actp.resize(h->ntiles + 1); ghp.resize(h->ntiles + 1);
    HIPCHK(h, hipMemcpy(h->d.active, actp.data(), actp.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d.ghosts, ghp.data(), ghp.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(h->d.eflags, 0, (size_t)h->ntiles * 4));
    std::vector<uint8_t> hs(h->ntiles, 0);
    // containsSource: tiles holding a cell within a flux source's radius
    // (src/SetSources.f90:357-375)
    for (int q = 0; q < h->cfg.n_sources; ++q) {
        double sx, sy, sr;
        hipMemcpy(&sx, h->d.src_x + q, 8, hipMemcpyDeviceToHost);
        hipMemcpy(&sy, h->d.src_y + q, 8, hipMemcpyDeviceToHost);
        hipMemcpy(&sr, h->d.src_r + q, 8, hipMemcpyDeviceToHost);
        for (int J = 0; J < h->NY; ++J)
            for (int I = 0; I < h->NX; ++I) {
                const int ti = I / h->nx, tj = J / h->ny, i = I % h->nx, j = J % h->ny;
                const double x = -0.5 * h->cfg.x_size + h->cfg.dx * ((ti + 1 - 1.0) * h->nx + (i + 1 - 0.5));
                const double y = -0.5 * h->cfg.y_size + h->cfg.dy * ((tj + 1 - 1.0) * h->ny + (j + 1 - 0.5));
                if ((x - sx) * (x - sx) + (y - sy) * (y - sy) <= sr * sr) hs[ti + tj * h->ntx] = 1;
            }
    }
    HIPCHK(h, hipMemcpy(h->d.has_src, hs.data(), hs.size(), hipMemcpyHostToDevice));
    Ctrl c;
    std::memset(&c, 0, sizeof c);
    c.t = t;
    c.tend = t;
    const unsigned long long huge = 0x7FEFFFFFFFFFFFFFull;
    c.slot[0] = c.slot[1] = c.slot[2] = huge;
    c.status = ST_DONE;
    c.n_active = (int)act.size();
    c.n_ghost = (int)gh.size();
    int rc = write_ctrl(h, &c);
    if (rc) return rc;
    if (!gh.empty())
        hipLaunchKernelGGL(init_ghosts_kernel, dim3(gh.size()), dim3(256), 0, h->stream, h->d, h->p);
    if (!act.empty())
        hipLaunchKernelGGL(init_flags_kernel, dim3(act.size()), dim3(256), 0, h->stream, h->d, h->p);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->t = t;
    h->have_state = true;
    return KST_OK;
}

static int start_run(kst_handle* h, int mode, double tend, long long nsteps, kst_stats* st) {
    if (!h) return KST_ERR_CONFIG;
    if (!h->have_state) return set_err(h, KST_ERR_STATE, "kst_set_state must come first");
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    if (c.status == ST_ERROR) return set_err(h, c.err, "handle is in an error state");
    c.status = ST_NEW;
    c.mode = mode;
    c.tend = tend;
    c.steps_left = nsteps;
    c.steps = c.rejected = c.rhs_evals = c.cell_updates = 0;
    c.min_dt = 0.0; c.max_dt = 0.0;
    rc = write_ctrl(h, &c);
    if (rc) return rc;
    const long long max_iters = mode == 1 ? 20 * nsteps + 1000 : 50000000LL;
    return run_until_done(h, st, max_iters);
}

int kst_integrate_to(kst_handle* h, double tend, kst_stats* st) {
    if (st) std::memset(st, 0, sizeof *st);
    if (!h) return KST_ERR_CONFIG;
    if (tend <= h->t) {
        if (st) st->t = h->t;
        return KST_OK;
    }
    return start_run(h, 0, tend, 0, st);
}

int kst_step_n(kst_handle* h, int64_t n, kst_stats* st) {
    if (st) std::memset(st, 0, sizeof *st);
    if (!h) return KST_ERR_CONFIG;
    if (n <= 0) return KST_OK;
    return start_run(h, 1, DBL_MAX, n, st);
}

int kst_get_state(kst_handle* h, double* w, double* rhu, double* rhv, double* hnpsi, double* hn,
                  double* btv, uint8_t* ts, double* t) {
    if (!h) return KST_ERR_CONFIG;
    const size_t FS = (size_t)h->ntiles * h->NC;
    std::vector<double> S(4 * FS);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(S.data(), h->d.S0, S.size() * 8, hipMemcpyDeviceToHost));
    std::vector<uint8_t> st(h->ntiles);
    HIPCHK(h, hipMemcpy(st.data(), h->d.tstate, st.size(), hipMemcpyDeviceToHost));
    // Fields outside allocated (active or ghost) tiles read as zero.
    for (int k = 0; k < h->ntiles; ++k)
        if (st[k] == KST_TILE_OFF)
            for (int q = 0; q < 4; ++q) std::fill_n(&S[q * FS + (size_t)k * h->NC], h->NC, 0.0);
    if (w) from_tiles(h, S.data(), w);
    if (rhu) from_tiles(h, S.data() + FS, rhu);
    if (rhv) from_tiles(h, S.data() + 2 * FS, rhv);
    if (hnpsi) from_tiles(h, S.data() + 3 * FS, hnpsi);
    if (hn) {
        std::vector<double> b0c(FS), g(FS), H(FS, 0.0);
        HIPCHK(h, hipMemcpy(b0c.data(), h->d.b0c, FS * 8, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(g.data(), h->d.cgam, FS * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < FS; ++i) {
            if (st[i / h->NC] == KST_TILE_OFF) continue;
            double x = -0.0;
            x = x + (S[i] - b0c[i]);
            x = x * g[i];
            H[i] = x < 0.0 ? 0.0 : x;
        }
        from_tiles(h, H.data(), hn);
    }
    if (btv) std::fill_n(btv, (size_t)(h->NX + 1) * (h->NY + 1), 0.0);
    if (ts) std::memcpy(ts, st.data(), st.size());
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    if (t) *t = c.t;
    return KST_OK;
}

int kst_get_maxima(kst_handle* h, double* const mx[11]) {
    if (!h || !mx) return KST_ERR_CONFIG;
    const size_t FS = (size_t)h->ntiles * h->NC;
    std::vector<double> M(11 * FS);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(M.data(), h->d.maxima, M.size() * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < 11; ++q)
        if (mx[q]) from_tiles(h, M.data() + q * FS, mx[q]);
    return KST_OK;
}

int kst_last_error(const kst_handle* h, char* buf, int buflen) {
    if (!h || !buf || buflen <= 0) return KST_ERR_CONFIG;
    std::snprintf(buf, buflen, "%s", h->err.c_str());
    return KST_OK;
}

int kst_profile(kst_handle* h, int enable) {
    if (!h) return KST_ERR_CONFIG;
    h->profiling = enable != 0;
    for (int k = 0; k < 4; ++k) { h->prof_ms[k] = 0.0; h->prof_n[k] = 0; }
    return KST_OK;
}

int kst_profile_read(kst_handle* h, double stage_ms[4], int64_t launches[4]) {
    if (!h) return KST_ERR_CONFIG;
    for (int k = 0; k < 4; ++k) {
        if (stage_ms) stage_ms[k] = h->prof_n[k] ? h->prof_ms[k] / h->prof_n[k] : 0.0;
        if (launches) launches[k] = h->prof_n[k];
        h->prof_ms[k] = 0.0;
        h->prof_n[k] = 0;
    }
    return KST_OK;
}

void kst_destroy(kst_handle* h) {
    if (!h) return;
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    if (h->stream) hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) hipFree(p);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
