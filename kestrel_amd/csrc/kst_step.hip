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
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/kestrel_step.h"
#include "kst_physics.hpp"

#pragma clang fp contract(off)

namespace kst {

constexpr int B = 16;            // sub-block edge (cells)
constexpr int RW = B + 4;        // region width with 2-cell halo
constexpr int NREG = RW * RW;
constexpr int NTHREADS = 256;
// Threads of a stage (and morphodynamic-stage) workgroup: 256 own the
// sub-block's cells.  With 512, 128 more would load the halo strips alongside
// them and all would share the faces; measured slower on MI355X (register
// pressure halves the resident waves), so the halo takes a second round.
constexpr int STAGE_NT = 256;

enum Status { ST_NEW = 0, ST_RETRY = 1, ST_DONE = 2, ST_ERROR = 3 };

// A CFL minimum (bit pattern of a positive double) on a cache line of its
// own: every block of a stage kernel updates one with an atomic, and the
// fields the blocks read at their start must not share its line (reads queue
// behind the atomics in L2).
struct alignas(128) CflSlot {
    unsigned long long bits;
};

struct Ctrl {
    double t, tend, dt_retry, last_dt, min_dt, max_dt;
    long long steps, rejected, rhs_evals, cell_updates;
    long long steps_left;
    long long redist_cells;    // cells corrected by RedistributeCell
    long long redist_ghost;    // of their writes, those into ghost tiles
    long long ndirty;          // ghost entries of intermed3 written in this step
    int status, err, n_active, n_ghost, mode;
    int mref;                  // morphodynamic step refused: 1 checks, 2 redistribution
    int ncand;                 // redistribution candidates of this attempt
    int commit;                // the step was accepted: the new bed becomes current
    long long work[16];        // launches that did work, per kernel (KW_*)
    int cur;                   // mixed precision: u^n is in S0 (0) or S0b (1)
    double next_t;             // NextFluxSeriesTime of t, kept with t
    // CFL minima: first hydraulic operator stages 1..3, then the second
    // operator of a Strang step stages 1..3
    CflSlot slot[6];
};

// Kernel ids of Ctrl::work.  Block 0 of a launch that does work counts it, so
// a profiler can tell productive launches from the no-ops that follow a
// rejected check or the end of a run.
enum {
    KW_STAGE = 0,      // + STAGE - 1, + 4 for the second hydraulic operator
    KW_MORPHO = 8,     // + S - 1
    KW_REDIST = 11,
    KW_FACETOPO = 12,
    KW_COMMIT = 13,
    KW_ACTIVATE = 14,
    KW_FINALIZE = 15,
    KW_N = 16
};

// An atomic whose result is unused: the thread does not wait for it (a
// read-modify-write would hold block 0 for a round trip before its barrier).
__device__ __forceinline__ void count_work(const Ctrl* c, int id) {
    if (blockIdx.x == 0) atomicAdd((unsigned long long*)&const_cast<Ctrl*>(c)->work[id], 1ULL);
}

// Bed-dependent topography of one generation, per tile.  Each tile holds its
// own copy of the vertices it shares with its neighbours (TileType%bt,
// src/Grid.f90:57-121); the copies may differ in the last bits because the
// reference computes their rates with different summation orders.
struct Topo {
    double* vbt;                               // vertex bt [tile][NV]
    double *btc, *cbx, *cby, *cgam;            // cell bt, db/dx, db/dy, gamma [tile][NC]
    double *pxu, *pxv, *pyv;                   // surface-gradient prefactors [tile][NC]
    double *xbt, *xby, *xgam, *xib;            // x-face bt, db/dy, gamma, CorrectSlopes bed [tile][NFX]
    double *ybt, *ybx, *ygam, *yib;            // y-face bt, db/dx, gamma, CorrectSlopes bed [tile][NFY]
    double *xnp, *xnm, *ynp, *ynm;             // CFL numerators gr^2*dx of the face's two cells
};

struct Dev {
    // ---- the hydraulic operator this launch evaluates
    const double *Uw, *Uu, *Uv, *Uh;  // its u^n: w, rhoHnu, rhoHnv, Hnpsi
    double* R;                        // RHS of its u^n: E_w, E_rhu, E_rhv, E_hnpsi, I
    double* S2;                       // stage 2
    double* S3;                       // stage 3
    double* Sout;                     // result of the final implicit pass
    int second;                       // 1: the second operator of a Strang step
    int last;                         // 1: ends the step (activation flags)
    Topo T;                           // the topography it sees
    // ---- step state and static data
    double* S0;                       // u^n of the step (tileContainer), 4 fields
    // Mixed precision keeps u^n in FP64 and the stages as FP32: R holds the
    // RHS of u^n, S2 and S3 the stages' increments over u^n, all float; the
    // last stage writes u^{n+1} to the other FP64 buffer (S0 and S0b swap
    // roles with Ctrl::cur when a step is accepted), since neighbouring
    // blocks still read u^n + S3 around their cells.
    double* S0b;
    int mixed;
    double *b0c, *xb0, *yb0, *vb0;    // b0 at cells, x-/y-faces, vertices [tile][...]
    double *gsx, *gsy;                // ghost w slopes [tile][NC]
    double* maxima;                   // [11][tile][NC]
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
    // ---- morphodynamics (src/TimeStepper.f90:532-781)
    double* SH;                       // u after the first hydraulic operator, 4 fields
    double* SM[3];                    // w, Hnpsi after morphodynamic stages 1..3
    double* UV;                       // velocities the morphodynamic RHS reads, 2 fields
    double* R2;                       // RHS of the second operator's u^n
    Topo TA;                          // bed at the start of the step (tileContainer)
    Topo TM[2];                       // bed after morphodynamic stages 1 and 2
    Topo TB;                          // bed after the morphodynamic operator
    long long* cand_key;              // redistribution candidates: scan-order key
    double* cand_ex;                  // and excess deposition
    int* cand_pos;                    // cell -> sorted candidate position (-1: none) [tile][NC]
    int cand_cap;
    int rd_chunk;                     // candidates walked at once (<= RD_THREADS)
    int rd_lds_sort;                  // largest padded candidate count sorted in LDS (<= RD_LDS)
    int act_cand;                     // activation candidates walked from LDS (<= MAXCAND)
    int morpho;
    // Redistribution may write ghost tiles' copies of shared vertices and
    // ghost cells' w, Hnpsi into intermed3 (src/Redistribute.f90:362-470); the
    // reference keeps them until the next step refreshes intermed3 from the
    // tileContainer.  Here they live in TB.vbt and S3 and are listed for
    // that refresh: bit 62 set = vertex tile*NV + v, else cell tile*NC + c.
    long long* dirty;
    int dirty_cap;
    // The velocities the RHS of the step's last stage derived, [2][tile][NC]:
    // tileContainer's u, v after CopySolutionData (src/TimeStepper.f90:811-835)
    // copies intermed3's u whole, i.e. before the final implicit drag update
    // of the momentum.  Written by stage 4 of the operator that ends the step;
    // tiles activated since hold 0, SetDefaultTileData's value.
    double* UVo;
};

// Kernel arguments (Dev, Params) that live in device memory.  Passed by value, every field
// a kernel reads is loaded into scalar registers at entry and held to the
// end, and the ~100 that do not fit are spilled; read through a pointer they
// are loaded where they are used.  The pointer is taken as one into the
// constant address space, so that the array pointers read from it are known
// to address global memory (global rather than flat loads and stores).
template <class T>
using ConstArg = const __attribute__((address_space(4))) T;

template <class T>
__device__ __forceinline__ ConstArg<T>& const_arg(const T* p) {
    return *(ConstArg<T>*)p;
}

// The same object as an ordinary reference (reads through it still become
// scalar loads once the compiler sees where it came from).
template <class T>
__device__ __forceinline__ const T& generic(ConstArg<T>& r) {
    return *(const T*)&r;
}

// ---------------------------------------------------------------------------
// time-step control shared by kernels
// ---------------------------------------------------------------------------

// NextFluxSeriesTime (src/TimeStepper.f90:281-305).
template <class DevT>
__device__ double next_flux_time(const DevT& d, const Params& p, double t) {
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

// The control block as kernels read it at their start: a pointer into the
// constant address space, so that the reads are scalar loads.  The fields
// read this way do not change while a kernel runs; its writes and atomics go
// through d.ctrl itself.
template <class DevT>
__device__ __forceinline__ ConstArg<Ctrl>* ctrl_view(const DevT& d) {
    return (ConstArg<Ctrl>*)d.ctrl;
}

// The fields of the control block that a kernel's start-of-launch checks
// read, loaded together: the CFL slots sit on cache lines of their own, and
// read behind each check's branch they would be fetched one after another.
// Has the member names of Ctrl, so that step_dt and first_operator_ok take
// it as they take the control block.
struct CtrlSnap {
    int status, n_active, mref, cur, ncand;
    double t, tend, next_t, dt_retry;
    struct {
        unsigned long long bits;
    } slot[6];
};

template <class CtrlP>
__device__ __forceinline__ CtrlSnap snap_ctrl(CtrlP c) {
    CtrlSnap s;
    s.status = c->status; s.n_active = c->n_active; s.mref = c->mref; s.cur = c->cur; s.ncand = c->ncand;
    s.t = c->t; s.tend = c->tend; s.next_t = c->next_t; s.dt_retry = c->dt_retry;
#pragma unroll
    for (int k = 0; k < 6; ++k) s.slot[k].bits = c->slot[k].bits;
    return s;
}

// The step's (hydraulic) dt: substep-1 advice clamped to the next
// output/flux-series time, or to half of it with morphodynamics
// (src/TimeStepper.f90:155-168), or the refined dt of a retried step.
// c: the control block or its constant view.
template <class DevT, class CtrlP>
__device__ double step_dt(const DevT& d, const Params& p, CtrlP c) {
    if (c->status == ST_RETRY) return c->dt_retry;
    double adv = 0.9 * advised_from_slot(p, c->slot[0].bits, c->n_active);
    double tmax = dmin(c->tend, c->next_t);
    if (!d.morpho) return dmin(adv, tmax - c->t);
    return dmin(adv, 0.5 * (tmax - c->t));
}

// The first hydraulic operator of the attempt passed both CFL re-checks
// (src/TimeStepper.f90:393,452).
template <class CtrlP>
__device__ __forceinline__ bool first_operator_ok(CtrlP c, const Params& p, double dt) {
    return advised_from_slot(p, c->slot[1].bits, c->n_active) >= dt &&
           advised_from_slot(p, c->slot[2].bits, c->n_active) >= dt;
}

// ---------------------------------------------------------------------------
// explicit source terms (src/Equations.f90:456-620)
// ---------------------------------------------------------------------------
template <class DevT>
__device__ void flux_sources(const DevT& d, const Params& p, double t, double grid_t, double x,
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
        if (p.oned) {   // per unit length on a 1D grid (src/Equations.f90:582-585)
            Qf = Qf / (double)d.src_cells[J] / p.dx;
            psiQf = psiQf / (double)d.src_cells[J] / p.dx;
        } else {
            Qf = Qf / (double)d.src_cells[J] / p.dx / p.dy;
            psiQf = psiQf / (double)d.src_cells[J] / p.dx / p.dy;
        }
        qsum = qsum + Qf;
        psum = psum + psiQf;
    }
    Qt = 0.0 + qsum;
    psiQt = 0.0 + psum;
}

// ---------------------------------------------------------------------------
// the fused stage kernel
// ---------------------------------------------------------------------------

// Limiters at compile time (src/Limiters.f90:83-184), or with LIM_RUNTIME the
// limiter rt a scenario of a batch carries.  MinMod1/2 use the identity
// max(-x,-y) = -min(x,y): negation and |.| are exact, so the value equals the
// reference's branchy form.
constexpr int LIM_RUNTIME = 100;

template <int LIM, typename T>
__device__ __forceinline__ T lim(int rt, T a, T b) {
    if (LIM == LIM_RUNTIME) {
        if (rt == KST_LIMITER_MINMOD2) return lim<KST_LIMITER_MINMOD2>(rt, a, b);
        if (rt == KST_LIMITER_MINMOD1) return lim<KST_LIMITER_MINMOD1>(rt, a, b);
        return (T)limiter(rt, (double)a, (double)b);
    } else if (LIM == KST_LIMITER_MINMOD2) {
        const T th = T(1.3);
        const T m = fmin(th * fabs(a), fmin(th * fabs(b), T(0.5) * fabs(a + b)));
        return (a * b <= T(0)) ? T(0) : copysign(m, a);
    } else if (LIM == KST_LIMITER_MINMOD1) {
        const T m = fmin(fabs(a), fabs(b));
        return (a * b <= T(0)) ? T(0) : copysign(m, a);
    } else {
        return (T)limiter(LIM, (double)a, (double)b);
    }
}

// Working set of one sub-block.  With MIX the per-cell Hnpsi, u, v and the
// fluxes other than the hydrostatic one (q = 4, in fx4/fy4) are float.
// BY: the sub-block's height (B or B/2); the region is RW x (BY + 4) cells.
// ONED: a 1D grid (one row of cells), which also keeps the conserved rhoHnv
// for its x-face reconstruction.
template <int NQ, bool MIX = false, int BY = B, bool ONED = false>
struct SmemT {
    using FT = typename std::conditional<MIX, float, double>::type;
    static constexpr int by = BY;
    static constexpr bool oned = ONED;
    static constexpr int NR = RW * (BY + 4);       // region cells
    static constexpr int NFX = (B + 1) * BY;       // x-faces
    static constexpr int NFY = B * (BY + 1);       // y-faces
    double w[NR];
    FT h[NR], u[NR], v[NR];
    double rho[NR];
    double rv[ONED ? NR : 1];  // 1D: the conserved rhoHnv
    double wst[NR];          // w behind the ghost tiles' stored face values (their tileContainer w)
    double xbed[BY][B + 3];  // CorrectSlopes bed of x-faces -1..B+1 of each row (relative)
    double ybed[B][BY + 3];  // CorrectSlopes bed of y-faces -1..BY+1 of each column
    double xside[2][BY];     // the W / E tile's own bed at this tile's edge x-faces
    double yside[2][B];      // the S / N tile's own bed at this tile's edge y-faces
    int gidx[NR];
    uint8_t act[NR];
    FT fx[NQ][NFX];          // hw, hru, hrv, hh, g (, pu, pv with eddy viscosity)
    FT fy[NQ][NFY];
    double fx4[MIX ? NFX : 1], fy4[MIX ? NFY : 1];
    double red[STAGE_NT / 64];
    int tile, bx0, by0, bxa, bya, skip, cur, more;
    double dt, tstage, tgrid;
};

__device__ __forceinline__ int rg(int rx, int ry) { return (ry + 2) * RW + (rx + 2); }

// Map a cell given relative to tile t (i in [-2, nx+1], j in [-2, ny+1], not
// both outside) to (tile, ci, cj).
__device__ __forceinline__ void map_cell(const Params& p, int t, int i, int j, int& tile, int& ci,
                                         int& cj) {
    ci = i; cj = j;
    int di = 0, dj = 0;
    if (i < 0) { di = -1; ci += p.nx; }
    else if (i >= p.nx) { di = 1; ci -= p.nx; }
    if (j < 0) { dj = -1; cj += p.ny; }
    else if (j >= p.ny) { dj = 1; cj -= p.ny; }
    tile = nb_tile(p, t, di, dj);
}

// The face topography a face's flux reads: db/d(along), gamma, b0 and bt of
// the face, and the CFL numerators of its two cells.
struct FaceTopo {
    double bp, g, b0, bt, np, nm;
};

template <bool YDIR, bool CFL, class DevT>
__device__ __forceinline__ FaceTopo load_face_topo_at(const DevT& d, const Params& p, int t, int bx0,
                                                      int by0, int fi, int fj) {
    const int fpos = (YDIR ? by0 : bx0) + fi;
    FaceTopo f;
    int fidx;
    if (!YDIR) {
        fidx = t * p.NFX + (by0 + fj) * (p.nx + 1) + fpos;
        f.bp = d.T.xby[fidx]; f.g = d.T.xgam[fidx]; f.b0 = d.xb0[fidx]; f.bt = d.T.xbt[fidx];
    } else {
        fidx = t * p.NFY + fpos * p.nx + (bx0 + fj);
        f.bp = d.T.ybx[fidx]; f.g = d.T.ygam[fidx]; f.b0 = d.yb0[fidx]; f.bt = d.T.ybt[fidx];
    }
    f.np = CFL ? (YDIR ? d.T.ynp : d.T.xnp)[fidx] : 0.0;
    f.nm = CFL ? (YDIR ? d.T.ynm : d.T.xnm)[fidx] : 0.0;
    return f;
}

template <bool YDIR, bool CFL, typename Smem, class DevT>
__device__ __forceinline__ FaceTopo load_face_topo(const DevT& d, const Params& p, const Smem& s,
                                                   int fi, int fj) {
    return load_face_topo_at<YDIR, CFL>(d, p, s.tile, s.bx0, s.by0, fi, fj);
}

// Fused work of one x- or y-face (CalculateLimitedDerivs + Reconstruct +
// CorrectSlopes + CalculateFluxes, src/HydraulicRHS.f90:181-1182).
// Along the face normal the four cells are c[-2], L=c[-1], R=c[0], c[1];
// fi is the face position along the normal, fj the row (x) / column (y).
template <bool YDIR, bool VISC, int LIM, bool CFL, typename Smem, class DevT>
__device__ __forceinline__ void face_flux(const DevT& d, const Params& p, Smem& s, int fi, int fj,
                                          double& cfl_min, const FaceTopo& ft) {
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
    const double sLw = actL ? dxr * lim<LIM>(p.limiter, wR - wL, wL - wLL) : gs[s.gidx[iL]];
    const double sRw = actR ? dxr * lim<LIM>(p.limiter, wRR - wR, wR - wL) : gs[s.gidx[iR]];
    const double hLL = s.h[iLL], hL = s.h[iL], hR = s.h[iR], hRR = s.h[iRR];
    const double sLh = actL ? dxr * lim<LIM>(p.limiter, hR - hL, hL - hLL) : 0.0;
    const double sRh = actR ? dxr * lim<LIM>(p.limiter, hRR - hR, hR - hL) : 0.0;
    const double uLL = s.u[iLL], uL = s.u[iL], uR = s.u[iR], uRR = s.u[iRR];
    const double sLu = actL ? dxr * lim<LIM>(p.limiter, uR - uL, uL - uLL) : 0.0;
    const double sRu = actR ? dxr * lim<LIM>(p.limiter, uRR - uR, uR - uL) : 0.0;
    const double vLL = s.v[iLL], vL = s.v[iL], vR = s.v[iR], vRR = s.v[iRR];
    const double sLv = actL ? dxr * lim<LIM>(p.limiter, vR - vL, vL - vLL) : 0.0;
    const double sRv = actR ? dxr * lim<LIM>(p.limiter, vRR - vR, vR - vL) : 0.0;
    const double rLL = s.rho[iLL], rL = s.rho[iL], rR = s.rho[iR], rRR = s.rho[iRR];
    const double sLr = actL ? dxr * lim<LIM>(p.limiter, rR - rL, rL - rLL) : 0.0;
    const double sRr = actR ? dxr * lim<LIM>(p.limiter, rRR - rR, rR - rL) : 0.0;

    // --- MUSCL face values (Reconstruct, src/HydraulicRHS.f90:397-490)
    const double wEL = wL + sLw * 0.5 * dx;   // L's east face
    const double wWR = wR - sRw * 0.5 * dx;   // R's west face

    // beds of the faces of L and R (InterpolateB, src/HydraulicRHS.f90:741-753);
    // the face itself is entry fi + 1, except at a tile edge, where the cell
    // on the far side sees its own tile's copy of the shared vertices
    const double* bed = YDIR ? &s.ybed[fj][0] : &s.xbed[fj][0];
    const double bWL = bed[fi], bER = bed[fi + 2];
    const double fbL = (fpos == 0) ? (YDIR ? s.yside[0][fj] : s.xside[0][fj]) : bed[fi + 1];
    const double fbR = (fpos == nn) ? (YDIR ? s.yside[1][fj] : s.xside[1][fj]) : bed[fi + 1];

    // --- CorrectSlopes for w (src/HydraulicRHS.f90:560-736): each cell's
    // pair of face values is replaced by the bed slope if either is below bed.
    // A ghost cell is tested with the face values ReconstructwAtEdges stored
    // from its tileContainer w (src/UpdateTiles.f90:669-750).
    const double wLc = s.wst[iL], wRc = s.wst[iR];
    double wFL = wEL, wFR = wWR;
    if ((wLc + sLw * 0.5 * dx < fbL) || (wLc - sLw * 0.5 * dx < bWL)) wFL = wL + 0.5 * (fbL - bWL);
    if ((wRc + sRw * 0.5 * dx < bER) || (wRc - sRw * 0.5 * dx < fbR)) wFR = wR + 0.5 * (fbR - bER);
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
    const double fbp = ft.bp, fg = ft.g, fb0 = ft.b0, fbt = ft.bt;
    // face Hn from the corrected w (src/HydraulicRHS.f90:492-517)
    const double HnL = compute_hn(wFL, fb0, fbt, fg);
    const double HnR = compute_hn(wFR, fb0, fbt, fg);
    const double unL = YDIR ? vFL : uFL, unR = YDIR ? vFR : uFR;
    const double ruL = rFL * HnL * uFL, ruR = rFR * HnR * uFR;
    double rvL, rvR;
    if constexpr (Smem::oned) {
        // a 1D grid keeps the x-face reconstruction of the conserved rhoHnv
        // (rho Hn v is rebuilt in 2D only, src/HydraulicRHS.f90:536-551)
        const double qLL = s.rv[iLL], qL = s.rv[iL], qR = s.rv[iR], qRR = s.rv[iRR];
        const double sLq = actL ? dxr * lim<LIM>(p.limiter, qR - qL, qL - qLL) : 0.0;
        const double sRq = actR ? dxr * lim<LIM>(p.limiter, qRR - qR, qR - qL) : 0.0;
        rvL = qL + sLq * 0.5 * dx;
        rvR = qR - sRq * 0.5 * dx;
    } else {
        rvL = rFL * HnL * vFL;
        rvR = rFR * HnR * vFR;
    }

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
        if (aPos > kEps) cfl_min = fmin(ft.np / aPos, cfl_min);
        if (fabs(aNeg) > kEps) cfl_min = fmin(ft.nm / fabs(aNeg), cfl_min);
    }

    // --- fluxes (src/HydraulicRHS.f90:1027-1067, src/Equations.f90:53-245)
    double* F = YDIR ? &s.fy[0][0] : &s.fx[0][0];
    constexpr int NFACE = YDIR ? Smem::NFY : Smem::NFX;
    const int fslot = YDIR ? (fi * B + fj) : (fj * (B + 1) + fi);
    // zero flux where the wave speeds close (src/HydraulicRHS.f90:1028-1031);
    // computed without a branch so the two faces of a thread interleave
    const double dif = aPos - aNeg;
    const bool none = dif < 1e-10;
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
    F[0 * NFACE + fslot] = none ? 0.0 : x / dif;
    x = ruR - ruL;
    x = x * aPos * aNeg;
    x = x + (aPos * cMu - aNeg * cPu);
    F[1 * NFACE + fslot] = none ? 0.0 : x / dif;
    x = rvR - rvL;
    x = x * aPos * aNeg;
    x = x + (aPos * cMv - aNeg * cPv);
    F[2 * NFACE + fslot] = none ? 0.0 : x / dif;
    x = hFR * fg - hFL * fg;
    x = x * aPos * aNeg;
    x = x + (aPos * cMh - aNeg * cPh);
    F[3 * NFACE + fslot] = none ? 0.0 : x / dif;
    F[4 * NFACE + fslot] = none ? 0.0 : (aPos * hydM - aNeg * hydP) / dif;
    if (VISC) {
        const double pPu = (HnR < 0.0) ? 0.0 : p.nu * rFR * HnR * sRu;
        const double pPv = (HnR < 0.0) ? 0.0 : p.nu * rFR * HnR * sRv;
        const double pMu = (HnL < 0.0) ? 0.0 : p.nu * rFL * HnL * sLu;
        const double pMv = (HnL < 0.0) ? 0.0 : p.nu * rFL * HnL * sLv;
        F[5 * NFACE + fslot] = none ? 0.0 : 0.5 * (pPu + pMu);
        F[6 * NFACE + fslot] = none ? 0.0 : 0.5 * (pPv + pMv);
    }
}

// face_flux in mixed precision.  The free surface stays FP64 end to end --
// w and its limited slopes, CorrectSlopes against the bed, the face Hn, the
// density and the hydrostatic flux 0.5 g rho hp^2 -- so a lake at rest
// balances its bed-slope source exactly as in FP64; Hnpsi, u, v, their
// slopes and face values, the wave speeds and the central-upwind
// combination of the other fluxes are float.
__device__ __forceinline__ float wave_root_f(const Params& p, float hn, float bperp, float gam) {
    if (hn <= 0.f) hn = 0.f;
    if (p.geometric) return __builtin_sqrtf((float)p.g * hn * (1.f + bperp * bperp) / (gam * gam * gam));
    return __builtin_sqrtf((float)p.g * hn);
}

template <bool YDIR, bool VISC, int LIM, bool CFL, typename Smem, class DevT>
__device__ __forceinline__ void face_flux_mixed(const DevT& d, const Params& p, Smem& s, int fi, int fj,
                                                double& cfl_min) {
    int iLL, iL, iR, iRR;
    if (!YDIR) { iLL = rg(fi - 2, fj); iL = rg(fi - 1, fj); iR = rg(fi, fj); iRR = rg(fi + 1, fj); }
    else       { iLL = rg(fj, fi - 2); iL = rg(fj, fi - 1); iR = rg(fj, fi); iRR = rg(fj, fi + 1); }
    const double dxr = YDIR ? p.dyr : p.dxr;
    const double dx = YDIR ? p.dy : p.dx;
    const float dxrf = (float)dxr, hdx = (float)(0.5 * dx);
    const bool actL = s.act[iL], actR = s.act[iR];
    const int nn = YDIR ? p.ny : p.nx;
    const int fpos = (YDIR ? s.by0 : s.bx0) + fi;

    // FP64: w slopes (non-active tiles keep their stored slope), density slopes
    const double wLL = s.w[iLL], wL = s.w[iL], wR = s.w[iR], wRR = s.w[iRR];
    const double* gs = YDIR ? d.gsy : d.gsx;
    const double sLw = actL ? dxr * lim<LIM>(p.limiter, wR - wL, wL - wLL) : gs[s.gidx[iL]];
    const double sRw = actR ? dxr * lim<LIM>(p.limiter, wRR - wR, wR - wL) : gs[s.gidx[iR]];
    const double rLL = s.rho[iLL], rL = s.rho[iL], rR = s.rho[iR], rRR = s.rho[iRR];
    const double sLr = actL ? dxr * lim<LIM>(p.limiter, rR - rL, rL - rLL) : 0.0;
    const double sRr = actR ? dxr * lim<LIM>(p.limiter, rRR - rR, rR - rL) : 0.0;
    // FP32: Hnpsi, u, v slopes
    const float hLL = s.h[iLL], hL = s.h[iL], hR = s.h[iR], hRR = s.h[iRR];
    const float sLh = actL ? dxrf * lim<LIM>(p.limiter, hR - hL, hL - hLL) : 0.f;
    const float sRh = actR ? dxrf * lim<LIM>(p.limiter, hRR - hR, hR - hL) : 0.f;
    const float uLL = s.u[iLL], uL = s.u[iL], uR = s.u[iR], uRR = s.u[iRR];
    const float sLu = actL ? dxrf * lim<LIM>(p.limiter, uR - uL, uL - uLL) : 0.f;
    const float sRu = actR ? dxrf * lim<LIM>(p.limiter, uRR - uR, uR - uL) : 0.f;
    const float vLL = s.v[iLL], vL = s.v[iL], vR = s.v[iR], vRR = s.v[iRR];
    const float sLv = actL ? dxrf * lim<LIM>(p.limiter, vR - vL, vL - vLL) : 0.f;
    const float sRv = actR ? dxrf * lim<LIM>(p.limiter, vRR - vR, vR - vL) : 0.f;

    // FP64: w face values, beds and CorrectSlopes (as face_flux)
    const double wEL = wL + sLw * 0.5 * dx;
    const double wWR = wR - sRw * 0.5 * dx;
    const double* bed = YDIR ? &s.ybed[fj][0] : &s.xbed[fj][0];
    const double bWL = bed[fi], bER = bed[fi + 2];
    const double fbL = (fpos == 0) ? (YDIR ? s.yside[0][fj] : s.xside[0][fj]) : bed[fi + 1];
    const double fbR = (fpos == nn) ? (YDIR ? s.yside[1][fj] : s.xside[1][fj]) : bed[fi + 1];
    const double wLc = s.wst[iL], wRc = s.wst[iR];
    double wFL = wEL, wFR = wWR;
    if ((wLc + sLw * 0.5 * dx < fbL) || (wLc - sLw * 0.5 * dx < bWL)) wFL = wL + 0.5 * (fbL - bWL);
    if ((wRc + sRw * 0.5 * dx < bER) || (wRc - sRw * 0.5 * dx < fbR)) wFR = wR + 0.5 * (fbR - bER);
    float hFL = hL + sLh * hdx, hFR = hR - sRh * hdx;
    if (fpos == 0) {
        if (hFL < 0.f) hFL = hL;
    } else {
        const float hWL = hL - sLh * hdx;
        if ((hFL < 0.f) || (hWL < 0.f)) hFL = hL;
    }
    if (fpos == nn) {
        if (hFR < 0.f) hFR = hR;
    } else {
        const float hER = hR + sRh * hdx;
        if ((hFR < 0.f) || (hER < 0.f)) hFR = hR;
    }
    const float uFL = uL + sLu * hdx, uFR = uR - sRu * hdx;
    const float vFL = vL + sLv * hdx, vFR = vR - sRv * hdx;
    const double rFL = rL + sLr * 0.5 * dx, rFR = rR - sRr * 0.5 * dx;

    const int t = s.tile;
    int fidx;
    double fbp, fg, fb0, fbt;
    if (!YDIR) {
        fidx = t * p.NFX + (s.by0 + fj) * (p.nx + 1) + fpos;
        fbp = d.T.xby[fidx]; fg = d.T.xgam[fidx]; fb0 = d.xb0[fidx]; fbt = d.T.xbt[fidx];
    } else {
        fidx = t * p.NFY + fpos * p.nx + (s.bx0 + fj);
        fbp = d.T.ybx[fidx]; fg = d.T.ygam[fidx]; fb0 = d.yb0[fidx]; fbt = d.T.ybt[fidx];
    }
    const double HnL = compute_hn(wFL, fb0, fbt, fg);
    const double HnR = compute_hn(wFR, fb0, fbt, fg);
    const float HnLf = (float)HnL, HnRf = (float)HnR, fgf = (float)fg;
    const float rFLf = (float)rFL, rFRf = (float)rFR;
    const float unL = YDIR ? vFL : uFL, unR = YDIR ? vFR : uFR;
    const float ruL = rFLf * HnLf * uFL, ruR = rFRf * HnRf * uFR;
    const float rvL = rFLf * HnLf * vFL, rvR = rFRf * HnRf * vFR;

    const float rP = wave_root_f(p, HnRf, (float)fbp, fgf);
    const float rM = wave_root_f(p, HnLf, (float)fbp, fgf);
    float wsP = unR + rP, wsM = unL + rM;
    float aPos = (wsP > wsM) ? wsP : wsM;
    if (aPos < 0.f) aPos = 0.f;
    wsP = unR - rP; wsM = unL - rM;
    float aNeg = (wsP < wsM) ? wsP : wsM;
    if (aNeg > 0.f) aNeg = 0.f;
    if (CFL) {
        if ((double)aPos > kEps)
            cfl_min = fmin((double)((float)(YDIR ? d.T.ynp : d.T.xnp)[fidx] / aPos), cfl_min);
        if ((double)fabsf(aNeg) > kEps)
            cfl_min = fmin((double)((float)(YDIR ? d.T.ynm : d.T.xnm)[fidx] / fabsf(aNeg)), cfl_min);
    }

    float* F = YDIR ? &s.fy[0][0] : &s.fx[0][0];
    constexpr int NFACE = YDIR ? Smem::NFY : Smem::NFX;
    double* F4 = YDIR ? s.fy4 : s.fx4;
    const int fslot = YDIR ? (fi * B + fj) : (fj * (B + 1) + fi);
    const float dif = aPos - aNeg;
    if (dif < 1e-10f) {
        for (int q = 0; q < (VISC ? 7 : 5); ++q) F[q * NFACE + fslot] = 0.f;
        F4[fslot] = 0.0;
        return;
    }
    const float rd = __builtin_amdgcn_rcpf(dif);
    const float cPw = HnRf * unR * fgf, cMw = HnLf * unL * fgf;
    const float cPh = hFR * unR * fgf, cMh = hFL * unL * fgf;
    const float cPu = ruR * unR, cMu = ruL * unL;
    const float cPv = rvR * unR, cMv = rvL * unL;
    const float ab = aPos * aNeg;
    F[0 * NFACE + fslot] = ((HnRf * fgf - HnLf * fgf) * ab + (aPos * cMw - aNeg * cPw)) * rd;
    F[1 * NFACE + fslot] = ((ruR - ruL) * ab + (aPos * cMu - aNeg * cPu)) * rd;
    F[2 * NFACE + fslot] = ((rvR - rvL) * ab + (aPos * cMv - aNeg * cPv)) * rd;
    F[3 * NFACE + fslot] = ((hFR * fgf - hFL * fgf) * ab + (aPos * cMh - aNeg * cPh)) * rd;
    // FP64 hydrostatic flux
    double hp = -fbt;
    hp = hp + (wFR - fb0);
    const double hydP = 0.5 * p.g * rFR * hp * hp;
    hp = -fbt;
    hp = hp + (wFL - fb0);
    const double hydM = 0.5 * p.g * rFL * hp * hp;
    const double aP = aPos, aN = aNeg;
    F4[fslot] = (aP * hydM - aN * hydP) / (aP - aN);
    if (VISC) {
        const float nu = (float)p.nu;
        float pPu, pPv, pMu, pMv;
        if (HnRf < 0.f) { pPu = 0.f; pPv = 0.f; }
        else { pPu = nu * rFRf * HnRf * sRu; pPv = nu * rFRf * HnRf * sRv; }
        if (HnLf < 0.f) { pMu = 0.f; pMv = 0.f; }
        else { pMu = nu * rFLf * HnLf * sLu; pMv = nu * rFLf * HnLf * sLv; }
        F[5 * NFACE + fslot] = 0.5f * (pPu + pMu);
        F[6 * NFACE + fslot] = 0.5f * (pPv + pMv);
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
        for (int i = 1; i < STAGE_NT / 64; ++i) m = fmin(m, s.red[i]);
        if (m < kHuge) atomicMin(slot, dbits(m));
    }
}

// InterpolateB (src/HydraulicRHS.f90:741-753) of tile tn's copy of vertices
// (i0, j0) and (i1, j1) in the bed the morphodynamic operator left (TB), which
// is intermed3's: a ghost tile's copy there may carry redistribution writes.
// On a 1D grid the bed of vertex (i0, j0) alone, b0 + bt (src/HydraulicRHS.f90:593-595).
template <bool ONED = false, class DevT>
__device__ __forceinline__ double ghost_bed(const DevT& d, const Params& p, int tn, int i0, int j0,
                                           int i1, int j1) {
    const size_t NV = (size_t)(p.nx + 1) * (p.ny + 1);
    const size_t a = (size_t)tn * NV + j0 * (p.nx + 1) + i0, b = (size_t)tn * NV + j1 * (p.nx + 1) + i1;
    if constexpr (ONED) return d.vb0[a] + d.TB.vbt[a];
    double ib = d.TB.vbt[a] + d.TB.vbt[b];
    ib = ib + d.vb0[a];
    ib = ib + d.vb0[b];
    return ib * 0.5;
}

// CorrectSlopes bed of a face given by its tile-relative index K in
// [-1, nn+1] along the normal and its row/column r, as tile t sees it
// (faces beyond the tile edge belong to the neighbour; they do not touch the
// shared vertices, so every tile sees the same value there).
template <bool YDIR, class DevT>
__device__ __forceinline__ double face_bed(const DevT& d, const Params& p, int t, int K, int r) {
    if (!YDIR) {
        if (K < 0) return d.T.xib[nb_tile(p, t, -1, 0) * p.NFX + r * (p.nx + 1) + (p.nx - 1)];
        if (K > p.nx) return d.T.xib[nb_tile(p, t, 1, 0) * p.NFX + r * (p.nx + 1) + 1];
        return d.T.xib[t * p.NFX + r * (p.nx + 1) + K];
    } else {
        if (K < 0) return d.T.yib[nb_tile(p, t, 0, -1) * p.NFY + (p.ny - 1) * p.nx + r];
        if (K > p.ny) return d.T.yib[nb_tile(p, t, 0, 1) * p.NFY + 1 * p.nx + r];
        return d.T.yib[t * p.NFY + K * p.nx + r];
    }
}

// STAGE 1: RHS(u^n) -> R.            (IntegrateTo, src/TimeStepper.f90:155,220)
// STAGE 2: u1 = f(u^n, R); RHS(u1); u2 -> S2.   (src/TimeStepper.f90:371-444)
// STAGE 3: RHS(u2); u3 -> S3.                   (src/TimeStepper.f90:447-498)
// STAGE 4: RHS(u3); implicit drag -> Sout; maxima; activation flags.
//                                               (src/TimeStepper.f90:501-524)
// pre: the face's topography when it was loaded ahead (FP64 only).
template <bool YDIR, bool VISC, int LIM, bool CFL, bool MIX, typename Smem, class DevT>
__device__ __forceinline__ void face_work(const DevT& d, const Params& p, Smem& s, int fi, int fj,
                                          double& cfl_min, const FaceTopo* pre = nullptr) {
    if constexpr (MIX) {
        face_flux_mixed<YDIR, VISC, LIM, CFL>(d, p, s, fi, fj, cfl_min);
    } else {
        const FaceTopo ft = pre ? *pre : load_face_topo<YDIR, CFL>(d, p, s, fi, fj);
        face_flux<YDIR, VISC, LIM, CFL>(d, p, s, fi, fj, cfl_min, ft);
    }
}

// Flux q of face slot i as FP64.
template <bool MIX, bool YDIR, typename Smem>
__device__ __forceinline__ double fluxv(const Smem& s, int q, int i) {
    if constexpr (MIX) {
        if (q == 4) return YDIR ? s.fy4[i] : s.fx4[i];
    }
    return YDIR ? (double)s.fy[q][i] : (double)s.fx[q][i];
}

// With morphodynamics the same kernels evaluate both hydraulic operators of
// the Strang step H(dt) M(2dt) H(dt); the second starts at t + dt from the
// morphodynamic result and first checks its substep-1 advice
// (src/TimeStepper.f90:214-258).
//
// The pass-th work item of this block in a one-scenario launch of G blocks
// (G a multiple of 8) over a list of nwork items.  Blocks are dealt to the 8
// XCDs round-robin, so XCD x (blocks x, x + 8, ...) takes the x-th contiguous
// eighth of the list and neighbouring sub-blocks share an L2; within it block
// b takes items b/8, b/8 + G/8, ...  Returns nwork when there is none; *more
// tells whether the block has a further item.
__device__ __forceinline__ int xcd_item(int nwork, int pass, bool* more) {
    const int G = gridDim.x, b = blockIdx.x, x = b % 8, m = G / 8;
    const int per = (nwork + 7) / 8;
    const int end = min((x + 1) * per, nwork);
    const int w = x * per + b / 8 + pass * m;
    *more = w + m < end;
    return w < end ? w : nwork;
}

// The small shape's block-to-item map, which needs no control-block read: XCD
// x (blocks x, x + 8, ...) takes items x G/8 .. x G/8 + G/8 - 1 of a launch of
// G blocks.  It covers the list while nwork <= G (the grid cap leaves room for
// the active set to double), and a block uses it only then, so that the load
// of its item's tile can be issued alongside the control-block reads.  (Large
// sets keep xcd_item: this map would leave XCDs idle when nwork < G.)
__device__ __forceinline__ int cap_item() {
    const int G = gridDim.x;
    return (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
}

// Item of thread tid when na items of one kind and nb of another (each at
// most 128) share a workgroup of four waves: the first kind on wave 0 and,
// beyond 64, on wave 3; the second on waves 1 and 2.  No wave mixes the two
// kinds (a mixed wave runs both code paths one after the other, so its
// neighbours wait at the next barrier for twice the work).  Returns the item
// index within its kind, or -1; *second tells the kind.
__device__ __forceinline__ int split_item(int tid, int na, int nb, bool* second) {
    if (tid < 64) { *second = false; return tid < na ? tid : -1; }
    if (tid < 192) { *second = true; const int g = tid - 64; return g < nb ? g : -1; }
    *second = false;
    const int f = tid - 128;   // 64 + (tid - 192)
    return f < na ? f : -1;
}

// stage_body does one work item (one sub-block of an active tile): witem is
// its index in the scenario's active-list order, or -1 to take the pass-th
// item of the block in a one-scenario launch.  Returns whether the block has
// a further item (false when it had none).
template <int STAGE, bool VISC, int LIM, bool MIX, typename Smem, class DevT>
__device__ __forceinline__ bool stage_body(const DevT& d, const Params& p, Smem& s, int witem,
                                           int pass = 0) {
    constexpr int BY = Smem::by;       // sub-block height
    // the thread index as a fresh value per pass: in a block's loop over
    // items nothing derived from it is hoisted and held across items (which
    // would cost registers, and resident waves, for the common single pass).
    // The asm has no side effects, so loads are still known unclobbered.
    unsigned tidu = threadIdx.x;
    asm("" : "+v"(tidu) : "s"(pass));
    const int tid = (int)(tidu & (STAGE_NT - 1));
    if (tid == 0) {
        const int sbx = (p.nx + B - 1) / B, sby = (p.ny + BY - 1) / BY, spt = sbx * sby;
        constexpr bool SPEC = BY < B;
        int wspec = 0, tspec = 0;
        if (SPEC && witem < 0) {
            wspec = cap_item();
            tspec = d.active[min(wspec / spt, p.ntiles)];   // the list holds ntiles + 1 entries
        }
        const CtrlSnap cs = snap_ctrl(ctrl_view(d));
        const CtrlSnap* c = &cs;
        int skip = 0;
        if (c->status == ST_DONE || c->status == ST_ERROR) skip = 1;
        if (STAGE == 1 && !d.second && c->status != ST_NEW) skip = 1;
        double dt = 0.0;
        if (!skip && (STAGE >= 2 || d.second)) {
            dt = step_dt(d, p, c);
            if (d.second && (!first_operator_ok(c, p, dt) || c->mref)) skip = 1;
            if (d.second && STAGE >= 2 && 0.9 * advised_from_slot(p, c->slot[3].bits, c->n_active) < dt) skip = 1;
            // constant slot indices keep the snapshot in registers
            const unsigned long long s1 = d.second ? c->slot[4].bits : c->slot[1].bits;
            const unsigned long long s2 = d.second ? c->slot[5].bits : c->slot[2].bits;
            if (STAGE >= 3 && advised_from_slot(p, s1, c->n_active) < dt) skip = 1;
            if (STAGE >= 4 && advised_from_slot(p, s2, c->n_active) < dt) skip = 1;
        }
        s.dt = dt;
        s.cur = c->cur;
        const double tb = d.second ? c->t + dt : c->t;   // grid%t of this operator
        s.tgrid = tb;
        s.tstage = (STAGE == 1) ? tb : (STAGE == 3 ? tb + 0.5 * dt : tb + dt);
        // work item -> (tile, sub-block); contiguous chunks of the active
        // list go to the same XCD (blocks are dealt round-robin to 8 XCDs)
        const int nwork = c->n_active * spt;
        const bool spec = SPEC && witem < 0 && nwork <= (int)gridDim.x;
        int w = witem;
        bool more = false;
        if (spec) w = pass == 0 ? wspec : nwork;
        else if (witem < 0) w = xcd_item(nwork, pass, &more);
        if (!skip && w == 0)
            atomicAdd((unsigned long long*)&d.ctrl->work[KW_STAGE + STAGE - 1 + (d.second ? 4 : 0)], 1ULL);
        if (w >= nwork) skip = 1;
        if (!skip) {
            s.tile = spec ? tspec : d.active[w / spt];
            const int sub = w % spt;
            s.bx0 = (sub % sbx) * B;
            s.by0 = (sub / sbx) * BY;
            s.bxa = min(B, p.nx - s.bx0);
            s.bya = min(BY, p.ny - s.by0);
        }
        s.skip = skip;
        s.more = more;
    }
    __syncthreads();
    // read as wave-uniform values, so that the block's loop over items is a
    // uniform loop (a loop the compiler must treat as divergent costs far more)
    if (__builtin_amdgcn_readfirstlane(s.skip)) return false;
    const bool more = __builtin_amdgcn_readfirstlane(s.more) != 0;
    const int t = s.tile;
    const double dt = s.dt;
    const int FS = (int)d.FS;
    const int bxa = s.bxa, bya = s.bya;

    // ---- phase A: stage state + derived variables on the region ----------
    const int tx = tid % B, ty = tid / B;
    const bool own = (tx < bxa) && (ty < bya);
    double ow = 0, oru = 0, orv = 0, oh = 0, ohn = 0, ou = 0, ov = 0, opsi = 0, orho = 0;
    double ob0 = 0, obt = 0, og = 1;
    double pw = 0, pru = 0, prv = 0, ph = 0;   // u^n of the owned cell (stage 2; mixed: 2..4)
    int oidx = 0;
    const double* Un = (MIX && s.cur) ? d.S0b : d.S0;   // u^n of the step
    // threads 0..OWN-1: the owned cells; the 2-cell halo strips (4 strips of
    // 2 x 16 slots) by the next 128 threads in the same round when there are
    // enough, else by threads 128..255 in a second round (waves 0 and 1 take
    // the extra faces of phase B instead)
    constexpr int OWN = B * BY;
    constexpr bool one_round = STAGE_NT >= OWN + 128;
    // A sub-block with at most one face per thread (the small shape) issues
    // its other global loads ahead of the state's so that they overlap: the
    // CorrectSlopes beds of its faces here (kept in a register until the
    // state is staged), below what phases B and C read.
    constexpr bool PRE = !MIX && Smem::NFX + Smem::NFY <= STAGE_NT;
    const bool pre = PRE && bxa == B && bya == BY;
    constexpr int NXB = BY * (B + 3), NYB = B * (BY + 3);   // x-face and y-face beds
    static_assert(!PRE || (NXB <= 128 && NYB <= 128 && Smem::NFX <= 128 && Smem::NFY <= 128),
                  "split_item: one bed and one face per thread, each kind within two waves");
    double bedv = 0.0;
    if (pre) {
        bool yb;
        const int q = split_item(tid, NXB, NYB, &yb);
        if (q >= 0) {
            if (!yb) bedv = face_bed<false>(d, p, t, s.bx0 + q % (B + 3) - 1, s.by0 + q / (B + 3));
            else bedv = face_bed<true>(d, p, t, s.by0 + q % (BY + 3) - 1, s.bx0 + q / (BY + 3));
        }
    }
    for (int round = 0; round < (one_round ? 1 : 2); ++round) {
        int rx, ry;
        bool valid;
        const int hq = one_round ? tid - OWN : tid - 128;
        if (round == 0 && tid < OWN) {
            rx = tx; ry = ty; valid = own;
        } else {
            if (hq < 0 || hq >= 128) break;
            const int strip = hq >> 5, k = hq & 31, dd = k >> 4, pos = k & 15;
            if (strip == 0) { rx = -2 + dd; ry = pos; valid = pos < bya; }
            else if (strip == 1) { rx = bxa + dd; ry = pos; valid = pos < bya; }
            else if (strip == 2) { rx = pos; ry = -2 + dd; valid = pos < bxa && !Smem::oned; }
            else { rx = pos; ry = bya + dd; valid = pos < bxa && !Smem::oned; }   // no rows beyond a 1D grid's
        }
        if (!valid) continue;
        int tc, ci, cj;
        map_cell(p, t, s.bx0 + rx, s.by0 + ry, tc, ci, cj);
        const int gi = tc * p.NC + cj * p.nx + ci;
        double w, ru, rv, h;
        if constexpr (MIX) {
            w = Un[gi]; ru = Un[FS + gi]; rv = Un[2 * FS + gi]; h = Un[3 * FS + gi];
            if (round == 0) { pw = w; pru = ru; prv = rv; ph = h; }
            if (STAGE == 2) {
                const float* Rf = (const float*)d.R;
                const double ew = Rf[gi], eu = Rf[FS + gi], ev = Rf[2 * FS + gi];
                const double eh = Rf[3 * FS + gi], im = Rf[4 * FS + gi];
                ru = (ru + dt * eu) / (1.0 - dt * im);
                rv = (rv + dt * ev) / (1.0 - dt * im);
                w = w + dt * ew;
                h = h + dt * eh;
            } else if (STAGE >= 3) {
                const float* Sf = (const float*)((STAGE == 3) ? d.S2 : d.S3);
                w = w + (double)Sf[gi];
                ru = ru + (double)Sf[FS + gi];
                rv = rv + (double)Sf[2 * FS + gi];
                h = h + (double)Sf[3 * FS + gi];
            }
        } else if (STAGE == 1 || STAGE == 2) {
            w = d.Uw[gi]; ru = d.Uu[gi]; rv = d.Uv[gi]; h = d.Uh[gi];
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
        const double b0c = d.b0c[gi], btc = d.T.btc[gi], gam = d.T.cgam[gi];
        const bool act = d.tstate[tc] == KST_TILE_ACTIVE;
        // ComputeDesingularisedVariables (src/HydraulicRHS.f90:762-878) runs
        // on active tiles only; ghost tiles keep SetDefaultTileData's zeros
        double hn = 0.0, psi = 0.0, rho = p.rhow, uu = 0.0, vv = 0.0;
        if (act) {
            hn = compute_hn(w, b0c, btc, gam);
            if (hn < 0.0) hn = 0.0;
            double hpc = h;
            if (hpc < 0.0) hpc = 0.0;
            const double eps = p.hthr;
            const double den = hn * hn + dmax(hn * hn, eps * eps);
            psi = dmin(2.0 * hn * hpc / den, p.max_pack);
            rho = density(p, psi);
            uu = 2.0 * hn * ru / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
            // a 1D grid never derives v (src/HydraulicRHS.f90:870-874): it stays 0
            if constexpr (!Smem::oned) vv = 2.0 * hn * rv / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
        } else if (p.bc == KST_BC_DIRICHLET && dirichlet_ghost(p, tc)) {
            hn = p.bc_hn; psi = p.bc_psi; rho = p.bc_rho; uu = p.bc_u; vv = p.bc_v;
        }
        const int r = rg(rx, ry);
        if constexpr (Smem::oned) s.rv[r] = rv;
        s.w[r] = w; s.h[r] = h; s.u[r] = uu; s.v[r] = vv; s.rho[r] = rho;
        s.wst[r] = act ? w : Un[gi];
        s.gidx[r] = gi;
        s.act[r] = act ? 1 : 0;
        if (round == 0) {
            ow = w; oru = ru; orv = rv; oh = h; ohn = hn; ou = uu; ov = vv; opsi = psi; orho = rho;
            ob0 = b0c; obt = btc; og = gam; oidx = gi;
        }
    }
    // the small shape's loads of what phases B and C read: its face's
    // topography, its cell's bed slopes and prefactors (and u^n in stage 3)
    // (the cell's bed slopes and prefactors also in the first and last stage
    // of the 16 x 16 shape, which have registers to spare for them)
    FaceTopo pft{};
    double pbx = 0, pby = 0, pgxu = 0, pgxv = 0, pgyv = 0, pw0 = 0, pu0 = 0, pv0 = 0, ph0 = 0;
    const bool prec = pre || STAGE == 1 || STAGE == 4;
    if (pre) {
        bool yf;
        const int f = split_item(tid, Smem::NFX, Smem::NFY, &yf);
        if (f >= 0) {
            if (!yf) pft = load_face_topo<false, STAGE != 4>(d, p, s, f % (B + 1), f / (B + 1));
            else pft = load_face_topo<true, STAGE != 4>(d, p, s, f / B, f % B);
        }
        if (own && STAGE == 3) { pw0 = d.Uw[oidx]; pu0 = d.Uu[oidx]; pv0 = d.Uv[oidx]; ph0 = d.Uh[oidx]; }
    }
    if (prec && own) {
        pbx = d.T.cbx[oidx]; pby = d.T.cby[oidx];
        pgxu = d.T.pxu[oidx]; pgxv = d.T.pxv[oidx]; pgyv = d.T.pyv[oidx];
    }
    // the last stage's maxima update reads u^n and six maxima fields of the
    // owned cell (each written only by its owner): loaded here too
    const bool trk = STAGE == 4 && p.track_maxima && !d.second;
    double mxv[6] = {0, 0, 0, 0, 0, 0}, mun[4] = {0, 0, 0, 0};
    if (trk && own) {
#pragma unroll
        for (int k = 0; k < 6; ++k) mxv[k] = d.maxima[2 * k * FS + oidx];   // fields 0, 2, ..., 10
#pragma unroll
        for (int k = 0; k < 4; ++k) mun[k] = Un[k * FS + oidx];
    }
    // face beds: rows of x-faces -1..bxa+1 and columns of y-faces -1..bya+1
    if (pre) {
        bool yb;
        const int q = split_item(tid, NXB, NYB, &yb);
        if (q >= 0) {
            if (!yb) s.xbed[q / (B + 3)][q % (B + 3)] = bedv;
            else s.ybed[q / (BY + 3)][q % (BY + 3)] = bedv;
        }
    } else {
        for (int q = tid; q < B * (B + 3); q += STAGE_NT) {
            const int r = q / (B + 3), k = q % (B + 3) - 1;
            if (r < bya && k <= bxa + 1) s.xbed[r][k + 1] = face_bed<false>(d, p, t, s.bx0 + k, s.by0 + r);
            if (!Smem::oned && r < bxa && k <= bya + 1)
                s.ybed[r][k + 1] = face_bed<true>(d, p, t, s.by0 + k, s.bx0 + r);
        }
    }
    // the neighbour tiles' own beds at the shared edge faces; in the last
    // stage a ghost neighbour's copy of the shared vertices is the one the
    // redistribution may have written (intermed3, src/Redistribute.f90:362-410)
    if (tid < B) {
        // the sides that apply, whether the neighbour is a ghost (in the last
        // stage of a morphodynamic step), then the loads issued together and
        // stored after: one round trip rather than one per side
        const int r = tid;
        const bool ovl = STAGE == 4 && d.morpho;
        const bool sx0 = r < bya && s.bx0 == 0, sx1 = r < bya && s.bx0 + bxa == p.nx;
        const bool sy0 = !Smem::oned && r < bxa && s.by0 == 0;
        const bool sy1 = !Smem::oned && r < bxa && s.by0 + bya == p.ny;
        const int row = (s.by0 + r) * (p.nx + 1), col = s.bx0 + r;
        const int tW = nb_tile(p, t, -1, 0), tE = nb_tile(p, t, 1, 0);
        const int tS = Smem::oned ? t : nb_tile(p, t, 0, -1), tN = Smem::oned ? t : nb_tile(p, t, 0, 1);
        const bool gW = ovl && sx0 && d.tstate[tW] != KST_TILE_ACTIVE;
        const bool gE = ovl && sx1 && d.tstate[tE] != KST_TILE_ACTIVE;
        const bool gS = ovl && sy0 && d.tstate[tS] != KST_TILE_ACTIVE;
        const bool gN = ovl && sy1 && d.tstate[tN] != KST_TILE_ACTIVE;
        double vW = 0.0, vE = 0.0, vS = 0.0, vN = 0.0;
        if (sx0)
            vW = gW ? ghost_bed<Smem::oned>(d, p, tW, p.nx, s.by0 + r, p.nx, s.by0 + r + 1)
                    : d.T.xib[tW * p.NFX + row + p.nx];
        if (sx1)
            vE = gE ? ghost_bed<Smem::oned>(d, p, tE, 0, s.by0 + r, 0, s.by0 + r + 1) : d.T.xib[tE * p.NFX + row];
        if (sy0)
            vS = gS ? ghost_bed(d, p, tS, col, p.ny, col + 1, p.ny) : d.T.yib[tS * p.NFY + p.ny * p.nx + col];
        if (sy1) vN = gN ? ghost_bed(d, p, tN, col, 0, col + 1, 0) : d.T.yib[tN * p.NFY + col];
        if (sx0) s.xside[0][r] = vW;
        if (sx1) s.xside[1][r] = vE;
        if (sy0) s.yside[0][r] = vS;
        if (sy1) s.yside[1][r] = vN;
    }
    __syncthreads();

    // ---- phase B: face fluxes --------------------------------------------
    double cflm = kHuge;
    // x-faces: consecutive lanes walk along a row; y-faces: along a row too
    if (bxa == B && bya == BY && BY == B && STAGE_NT == 256) {
        // 272 x-faces and 272 y-faces: one of each per thread, the 16 extra
        // x-faces on wave 0 and the 16 extra y-faces on wave 1
        face_work<false, VISC, LIM, STAGE != 4, MIX>(d, p, s, tid % (B + 1), tid / (B + 1), cflm);
        face_work<true, VISC, LIM, STAGE != 4, MIX>(d, p, s, tid / B, tid % B, cflm);
        if (tid < 16) {
            const int f = 256 + tid;
            face_work<false, VISC, LIM, STAGE != 4, MIX>(d, p, s, f % (B + 1), f / (B + 1), cflm);
        } else if (tid >= 64 && tid < 80) {
            const int f = 256 + tid - 64;
            face_work<true, VISC, LIM, STAGE != 4, MIX>(d, p, s, f / B, f % B, cflm);
        }
    } else if (pre) {
        bool yf;
        const int f = split_item(tid, Smem::NFX, Smem::NFY, &yf);
        if (f >= 0) {
            if (!yf) face_work<false, VISC, LIM, STAGE != 4, MIX>(d, p, s, f % (B + 1), f / (B + 1), cflm, &pft);
            else face_work<true, VISC, LIM, STAGE != 4, MIX>(d, p, s, f / B, f % B, cflm, &pft);
        }
    } else if (bxa == B && bya == BY) {
        // x- and y-faces as one list over the block
        for (int f = tid; f < Smem::NFX + Smem::NFY; f += STAGE_NT) {
            if (f < Smem::NFX) {
                face_work<false, VISC, LIM, STAGE != 4, MIX>(d, p, s, f % (B + 1), f / (B + 1), cflm);
            } else {
                const int g = f - Smem::NFX;
                face_work<true, VISC, LIM, STAGE != 4, MIX>(d, p, s, g / B, g % B, cflm);
            }
        }
    } else {
        for (int f = tid; f < (bxa + 1) * bya; f += STAGE_NT)
            face_work<false, VISC, LIM, STAGE != 4, MIX>(d, p, s, f % (bxa + 1), f / (bxa + 1), cflm);
        if constexpr (!Smem::oned)   // no y faces on a 1D grid (src/HydraulicRHS.f90:1075)
            for (int f = tid; f < (bya + 1) * bxa; f += STAGE_NT)
                face_work<true, VISC, LIM, STAGE != 4, MIX>(d, p, s, f / bxa, f % bxa, cflm);
    }
    if (STAGE != 4) {
        reduce_min_to(s, cflm, &d.ctrl->slot[(d.second ? 3 : 0) + STAGE - 1].bits);
    } else {
        __syncthreads();
    }

    // ---- phase C: RHS of the owned cell and the stage update -------------
    if (!own) return more;
    const int xs = ty * (B + 1) + tx, ys = ty * B + tx;
#define FX(q, o) fluxv<MIX, false>(s, q, xs + (o))
#define FY(q, o) fluxv<MIX, true>(s, q, ys + (o))
    const double gam = og;
    const double bx = prec ? pbx : d.T.cbx[oidx], by = prec ? pby : d.T.cby[oidx];
    // prefactors (1+by^2)/gam, -(bx by/gam), (1+bx^2)/gam are per-bed topography
    const double gxu = prec ? pgxu : d.T.pxu[oidx], gxv = prec ? pgxv : d.T.pxv[oidx], gyu = gxv;
    const double gyv = prec ? pgyv : d.T.pyv[oidx];
    // ConstructHydraulicRHS (src/HydraulicRHS.f90:1228-1271; on a 1D grid
    // :1276-1294, x fluxes only and the hydrostatic term divided by gamma)
    double stw, sth, stu, stv;
    if constexpr (Smem::oned) {
        stw = (FX(0, 0) - FX(0, 1)) * p.dxr / (gam * gam);
        sth = (FX(3, 0) - FX(3, 1)) * p.dxr / gam;
        Kahan k1;
        k1.add(FX(1, 0) - FX(1, 1));
        k1.add((FX(4, 0) - FX(4, 1)) / gam);
        k1.add(VISC ? (FX(5, 1) - FX(5, 0)) : 0.0);
        stu = k1.s * p.dxr;
        Kahan k3;
        k3.add(FX(2, 0) - FX(2, 1));
        k3.add((FX(4, 0) - FX(4, 1)) / gam);
        k3.add(VISC ? (FX(6, 1) - FX(6, 0)) : 0.0);
        stv = k3.s * p.dxr;
    } else {
        stw = (FX(0, 0) - FX(0, 1)) * p.dxr / (gam * gam) + (FY(0, 0) - FY(0, B)) * p.dyr / (gam * gam);
        sth = (FX(3, 0) - FX(3, 1)) * p.dxr / gam + (FY(3, 0) - FY(3, B)) * p.dyr / gam;
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
        flux_sources(d, p, s.tstage, s.tgrid, x, y, Qt, psiQt);
        Ew = stw + (0.0 + Qt / (gam * gam));
        Eh = sth + (0.0 + psiQt / gam);
    } else {
        Ew = stw + 0.0;
        Eh = sth + 0.0;
    }
    double hpg = -obt;
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
        if constexpr (MIX) {
            float* Rf = (float*)d.R;
            Rf[oidx] = (float)Ew; Rf[FS + oidx] = (float)Eu; Rf[2 * FS + oidx] = (float)Ev;
            Rf[3 * FS + oidx] = (float)Eh; Rf[4 * FS + oidx] = (float)Im;
        } else {
            d.R[oidx] = Ew; d.R[FS + oidx] = Eu; d.R[2 * FS + oidx] = Ev; d.R[3 * FS + oidx] = Eh;
            d.R[4 * FS + oidx] = Im;
        }
    } else if (STAGE == 2) {
        // second SSP substep (src/TimeStepper.f90:407-444)
        const double u2 = 0.75 * pru + 0.25 * (oru + dt * Eu) / (1.0 - dt * Im);
        const double v2 = 0.75 * prv + 0.25 * (orv + dt * Ev) / (1.0 - dt * Im);
        double hpo = -obt;
        hpo = hpo + (pw - ob0);
        double hpn = -obt;
        hpn = hpn + (ow - ob0);
        double wu = obt;
        wu = wu + 0.25 * hpn;
        wu = wu + 0.75 * hpo;
        wu = wu + 0.25 * dt * Ew;
        wu = wu + ob0;
        const double h2 = 0.75 * ph + 0.25 * (oh + dt * Eh);
        if constexpr (MIX) {
            float* Sf = (float*)d.S2;
            Sf[oidx] = (float)(wu - pw); Sf[FS + oidx] = (float)(u2 - pru);
            Sf[2 * FS + oidx] = (float)(v2 - prv); Sf[3 * FS + oidx] = (float)(h2 - ph);
        } else {
            d.S2[oidx] = wu; d.S2[FS + oidx] = u2; d.S2[2 * FS + oidx] = v2; d.S2[3 * FS + oidx] = h2;
        }
    } else if (STAGE == 3) {
        // third SSP substep (src/TimeStepper.f90:466-498)
        const double w0 = MIX ? pw : (pre ? pw0 : d.Uw[oidx]), u0 = MIX ? pru : (pre ? pu0 : d.Uu[oidx]);
        const double v0 = MIX ? prv : (pre ? pv0 : d.Uv[oidx]), h0 = MIX ? ph : (pre ? ph0 : d.Uh[oidx]);
        const double third = 1.0 / 3.0, twothird = 2.0 / 3.0;
        const double u3 = third * u0 + twothird * (oru + dt * Eu) / (1.0 - dt * Im);
        const double v3 = third * v0 + twothird * (orv + dt * Ev) / (1.0 - dt * Im);
        double hpo = -obt;
        hpo = hpo + (w0 - ob0);
        double hpn = -obt;
        hpn = hpn + (ow - ob0);
        double wu = obt;
        wu = wu + third * hpo;
        wu = wu + twothird * hpn;
        wu = wu + twothird * dt * Ew;
        wu = wu + ob0;
        const double h3 = third * h0 + twothird * (oh + dt * Eh);
        if constexpr (MIX) {
            float* Sf = (float*)d.S3;
            Sf[oidx] = (float)(wu - w0); Sf[FS + oidx] = (float)(u3 - u0);
            Sf[2 * FS + oidx] = (float)(v3 - v0); Sf[3 * FS + oidx] = (float)(h3 - h0);
        } else {
            d.S3[oidx] = wu; d.S3[FS + oidx] = u3; d.S3[2 * FS + oidx] = v3; d.S3[3 * FS + oidx] = h3;
        }
    } else {
        // final implicit substep (src/TimeStepper.f90:504-518)
        const double un = (oru - dt * dt * Eu * Im) / (1.0 + dt * dt * Im * Im);
        const double vn = (orv - dt * dt * Ev * Im) / (1.0 + dt * dt * Im * Im);
        if (trk) {
            // UpdateMaximum* on the pre-step state at nextT (src/TimeStepper.f90:1155-1303);
            // the second operator of a Strang step repeats the first one's
            // update with the same state, which changes nothing
            const double w0 = mun[0], u0 = mun[1], v0 = mun[2], h0 = mun[3];
            const double tn = s.tgrid + dt;       // the step's t + dt
            double* M = d.maxima;
            if (obt < 0.0 && -obt > mxv[2]) { M[4 * FS + oidx] = -obt; M[5 * FS + oidx] = tn; }
            if (obt > 0.0 && obt > mxv[3]) { M[6 * FS + oidx] = obt; M[7 * FS + oidx] = tn; }
            double hn = compute_hn(w0, ob0, obt, gam);
            if (hn < 0.0) hn = 0.0;
            if (hn > p.hthr) {
                double hp0 = h0;
                if (hp0 < 0.0) hp0 = 0.0;
                const double eps = p.hthr;
                const double den = hn * hn + dmax(hn * hn, eps * eps);
                const double psi = dmin(2.0 * hn * hp0 / den, p.max_pack);
                const double rho = density(p, psi);
                const double uu = 2.0 * hn * u0 / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
                const double vv = Smem::oned ? 0.0 : 2.0 * hn * v0 / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
                if (mxv[5] == -1) M[10 * FS + oidx] = tn;
                if (hn > mxv[0]) { M[oidx] = hn; M[FS + oidx] = tn; }
                const double spd = sqrt(speed2(p.geometric, uu, vv, bx, by));
                if (spd > mxv[1]) { M[2 * FS + oidx] = spd; M[3 * FS + oidx] = tn; }
                if (psi > mxv[4]) { M[8 * FS + oidx] = psi; M[9 * FS + oidx] = tn; }
            }
        }
        double* Out = MIX ? (s.cur ? d.S0 : d.S0b) : d.Sout;
        Out[oidx] = ow; Out[FS + oidx] = un; Out[2 * FS + oidx] = vn; Out[3 * FS + oidx] = oh;
        if (d.last) { d.UVo[oidx] = ou; d.UVo[FS + oidx] = ov; }
        // activation flags from this state's Hn (CheckIfNearBoundaries,
        // src/TimeStepper.f90:946-1150): bit 0 N, 1 S, 2 E, 3 W
        if (d.last && ohn > p.hthr) {
            const int i = s.bx0 + tx, j = s.by0 + ty, tb = p.tile_buffer;
            unsigned f = 0;
            if (j + 1 > p.ny - tb) f |= 1u;
            if (j + 1 <= tb) f |= 2u;
            if (i + 1 > p.nx - tb) f |= 4u;
            if (i + 1 <= tb) f |= 8u;
            if (f) atomicOr(&d.eflags[t], f);
        }
    }
    return more;
}

// The grid is sized for the active set the host last saw (with room to
// grow); a block whose share of the list holds more items takes them in turn.
template <int STAGE, bool VISC, int LIM, bool MIX = false, int BY = B, bool ONED = false>
__global__ __launch_bounds__(STAGE_NT) void stage_kernel(const Dev* __restrict__ d,
                                                         const Params* __restrict__ p) {
    __shared__ SmemT<VISC ? 7 : 5, MIX, BY, ONED> s;
    for (int pass = 0; stage_body<STAGE, VISC, LIM, MIX>(const_arg(d), *p, s, -1, pass); ++pass)
        __syncthreads();
}

// ---------------------------------------------------------------------------
// topography
// ---------------------------------------------------------------------------

// Cell and face topography of tile t from its vertex elevations
// (ComputeCellCentredTopographicData / ComputeInterfacialTopographicData,
// src/MorphodynamicRHS.f90:308-578; face b0 of GetHeights, src/dem.f90:378-386;
// CorrectSlopes beds of InterpolateB, src/HydraulicRHS.f90:741-753).  Block-wide.
#define V(a, i, j) a[(j) * NV1 + (i)]
template <class DevT, class TopoT>
__device__ void tile_cell_topo(const DevT& d, const Params& p, const TopoT& T, int t, bool zero_btc) {
    const int nx = p.nx, NV1 = nx + 1;
    const size_t NV = (size_t)(nx + 1) * (p.ny + 1);
    const double* b = d.vb0 + (size_t)t * NV;
    const double* bt = T.vbt + (size_t)t * NV;
    if (p.oned) {
        // 1D: two-vertex forms along the lower vertex row (src/MorphodynamicRHS.f90:348-357)
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const double b0c = 0.5 * (b[i] + b[i + 1]);
            const double btc = zero_btc ? 0.0 : 0.5 * (bt[i] + bt[i + 1]);
            Kahan kx;
            kx.add(b[i + 1]); kx.add(bt[i + 1]); kx.add(-b[i]); kx.add(-bt[i]);
            const double dbdx = p.dxr * kx.s;
            const size_t gi = (size_t)t * p.NC + i;
            d.b0c[gi] = b0c; T.btc[gi] = btc; T.cbx[gi] = dbdx; T.cby[gi] = 0.0;
            T.cgam[gi] = gamma_of(p.geometric, dbdx, 0.0);
        }
        return;
    }
    for (int c = threadIdx.x; c < p.NC; c += blockDim.x) {
        const int i = c % nx, j = c / nx;
        Kahan k;
        k.add(V(b, i, j)); k.add(V(b, i + 1, j)); k.add(V(b, i, j + 1)); k.add(V(b, i + 1, j + 1));
        const double b0c = 0.25 * k.s;
        Kahan kt;
        kt.add(V(bt, i, j)); kt.add(V(bt, i + 1, j)); kt.add(V(bt, i, j + 1)); kt.add(V(bt, i + 1, j + 1));
        const double btc = zero_btc ? 0.0 : 0.25 * kt.s;
        Kahan kx;
        kx.add(V(b, i + 1, j)); kx.add(V(bt, i + 1, j)); kx.add(-V(b, i, j)); kx.add(-V(bt, i, j));
        kx.add(V(b, i + 1, j + 1)); kx.add(V(bt, i + 1, j + 1)); kx.add(-V(b, i, j + 1)); kx.add(-V(bt, i, j + 1));
        const double dbdx = 0.5 * p.dxr * kx.s;
        Kahan ky;
        ky.add(V(b, i, j + 1)); ky.add(V(bt, i, j + 1)); ky.add(-V(b, i, j)); ky.add(-V(bt, i, j));
        ky.add(V(b, i + 1, j + 1)); ky.add(V(bt, i + 1, j + 1)); ky.add(-V(b, i + 1, j)); ky.add(-V(bt, i + 1, j));
        const double dbdy = 0.5 * p.dyr * ky.s;
        const size_t gi = (size_t)t * p.NC + c;
        d.b0c[gi] = b0c; T.btc[gi] = btc; T.cbx[gi] = dbdx; T.cby[gi] = dbdy;
        T.cgam[gi] = gamma_of(p.geometric, dbdx, dbdy);
    }
}

// The W, E, S, N neighbours of tile t; without periodic boundaries a tile on
// the domain edge is its own neighbour beyond the edge.
__device__ __forceinline__ void side_tiles(const Params& p, int t, int& tW, int& tE, int& tS, int& tN) {
    if (p.periodic) {
        tW = nb_tile(p, t, -1, 0); tE = nb_tile(p, t, 1, 0);
        tS = nb_tile(p, t, 0, -1); tN = nb_tile(p, t, 0, 1);
        return;
    }
    const int ti = t % p.ntx, tj = t / p.ntx;
    tW = ti > 0 ? t - 1 : t; tE = ti < p.ntx - 1 ? t + 1 : t;
    tS = tj > 0 ? t - p.ntx : t; tN = tj < p.nty - 1 ? t + p.ntx : t;
}

// part/nparts: the share of the tile's faces one block of several computes.
// With STATIC the faces' CFL numerators (tile_static_terms) are computed in
// the same pass; they read the cell gamma of the tiles on both sides.
template <bool STATIC = false, class DevT, class TopoT>
__device__ void tile_face_topo(const DevT& d, const Params& p, const TopoT& T, int t, int part = 0,
                               int nparts = 1) {
    const int lane0 = threadIdx.x + part * blockDim.x, stride = blockDim.x * nparts;
    const int nx = p.nx, ny = p.ny, NV1 = nx + 1;
    // neighbour tiles (without periodic boundaries clamped at the domain
    // edge, where their values are never read)
    int tW, tE, tS, tN;
    side_tiles(p, t, tW, tE, tS, tN);
    const size_t NV = (size_t)(nx + 1) * (ny + 1);
    const double* b = d.vb0 + (size_t)t * NV;
    const double* bt = T.vbt + (size_t)t * NV;
    const double* bW = d.vb0 + (size_t)tW * NV; const double* btW = T.vbt + (size_t)tW * NV;
    const double* bE = d.vb0 + (size_t)tE * NV; const double* btE = T.vbt + (size_t)tE * NV;
    const double* bS = d.vb0 + (size_t)tS * NV; const double* btS = T.vbt + (size_t)tS * NV;
    const double* bN = d.vb0 + (size_t)tN * NV; const double* btN = T.vbt + (size_t)tN * NV;
    if (p.oned) {
        // 1D (src/MorphodynamicRHS.f90:545-577, src/dem.f90:387-390,
        // src/HydraulicRHS.f90:593-595): x faces of the lower vertex row, the
        // face's own vertex b0 and bt, a two-vertex central dbdx, no
        // perpendicular slope, the CorrectSlopes bed b0 + bt
        for (int i = lane0; i <= nx; i += stride) {
            double gl = 0.0, gr_ = 0.0;
            if (STATIC) {
                gl = (i == 0) ? T.cgam[tW * p.NC + nx - 1] : T.cgam[t * p.NC + i - 1];
                gr_ = (i == nx) ? T.cgam[tE * p.NC] : T.cgam[t * p.NC + i];
            }
            Kahan kx;
            if (i >= 1 && i <= nx - 1) {
                kx.add(V(b, i + 1, 0)); kx.add(V(bt, i + 1, 0)); kx.add(-V(b, i - 1, 0)); kx.add(-V(bt, i - 1, 0));
            } else if (i == 0) {
                kx.add(V(b, 1, 0)); kx.add(V(bt, 1, 0)); kx.add(-V(bW, nx - 1, 0)); kx.add(-V(btW, nx - 1, 0));
            } else {
                kx.add(V(bE, 1, 0)); kx.add(V(btE, 1, 0)); kx.add(-V(b, nx - 1, 0)); kx.add(-V(bt, nx - 1, 0));
            }
            const double dbdx = 0.5 * p.dxr * kx.s;
            const size_t fi = (size_t)t * p.NFX + i;
            const double fg = gamma_of(p.geometric, dbdx, 0.0);
            d.xb0[fi] = V(b, i, 0);
            T.xbt[fi] = V(bt, i, 0);
            T.xby[fi] = 0.0;
            T.xgam[fi] = fg;
            T.xib[fi] = V(b, i, 0) + V(bt, i, 0);
            if (STATIC) {
                const double a = fmin(gl / fg, 1.0), c = fmin(gr_ / fg, 1.0);
                T.xnp[fi] = a * a * p.dx;
                T.xnm[fi] = c * c * p.dx;
            }
        }
        return;
    }
    // x-faces: i in [0, nx], j in [0, ny)
    for (int f = lane0; f < (nx + 1) * ny; f += stride) {
        const int i = f % (nx + 1), j = f / (nx + 1);
        double gl = 0.0, gr_ = 0.0;
        if (STATIC) {
            gl = (i == 0) ? T.cgam[tW * p.NC + j * nx + nx - 1] : T.cgam[t * p.NC + j * nx + i - 1];
            gr_ = (i == nx) ? T.cgam[tE * p.NC + j * nx] : T.cgam[t * p.NC + j * nx + i];
        }
        // the vertex columns either side: this tile's, or across the tile
        // edge the neighbour's (selected per lane, so that a wave's edge and
        // interior faces load together rather than one branch after another)
        const double* bR = (i == nx) ? bE : b;
        const double* btR = (i == nx) ? btE : bt;
        const int iR = (i == nx) ? 1 : i + 1;
        const double* bL = (i == 0) ? bW : b;
        const double* btL = (i == 0) ? btW : bt;
        const int iL = (i == 0) ? nx - 1 : i - 1;
        Kahan kx;
        kx.add(V(bR, iR, j)); kx.add(V(btR, iR, j)); kx.add(V(bR, iR, j + 1)); kx.add(V(btR, iR, j + 1));
        kx.add(-V(bL, iL, j)); kx.add(-V(btL, iL, j)); kx.add(-V(bL, iL, j + 1)); kx.add(-V(btL, iL, j + 1));
        const double dbdx = 0.25 * p.dxr * kx.s;
        const double btf = 0.5 * (V(bt, i, j) + V(bt, i, j + 1));
        Kahan ky;
        ky.add(V(b, i, j + 1)); ky.add(V(bt, i, j + 1)); ky.add(-V(b, i, j)); ky.add(-V(bt, i, j));
        const double dbdy = p.dyr * ky.s;
        double ib = V(bt, i, j) + V(bt, i, j + 1);
        ib = ib + V(b, i, j);
        ib = ib + V(b, i, j + 1);
        ib = ib * 0.5;
        const size_t fi = (size_t)t * p.NFX + f;
        const double fg = gamma_of(p.geometric, dbdx, dbdy);
        d.xb0[fi] = 0.5 * (V(b, i, j) + V(b, i, j + 1));
        T.xbt[fi] = btf;
        T.xby[fi] = dbdy;
        T.xgam[fi] = fg;
        T.xib[fi] = ib;
        if (STATIC) {
            const double a = fmin(gl / fg, 1.0), c = fmin(gr_ / fg, 1.0);
            T.xnp[fi] = a * a * p.dx;
            T.xnm[fi] = c * c * p.dx;
        }
    }
    // y-faces: i in [0, nx), j in [0, ny]
    for (int f = lane0; f < nx * (ny + 1); f += stride) {
        const int i = f % nx, j = f / nx;
        double gl = 0.0, gr_ = 0.0;
        if (STATIC) {
            gl = (j == 0) ? T.cgam[tS * p.NC + (ny - 1) * nx + i] : T.cgam[t * p.NC + (j - 1) * nx + i];
            gr_ = (j == ny) ? T.cgam[tN * p.NC + i] : T.cgam[t * p.NC + j * nx + i];
        }
        Kahan kx;
        kx.add(V(b, i + 1, j)); kx.add(V(bt, i + 1, j)); kx.add(-V(b, i, j)); kx.add(-V(bt, i, j));
        const double dbdx = p.dxr * kx.s;
        const double btf = 0.5 * (V(bt, i, j) + V(bt, i + 1, j));
        const double* bU = (j == ny) ? bN : b;
        const double* btU = (j == ny) ? btN : bt;
        const int jU = (j == ny) ? 1 : j + 1;
        const double* bD = (j == 0) ? bS : b;
        const double* btD = (j == 0) ? btS : bt;
        const int jD = (j == 0) ? ny - 1 : j - 1;
        Kahan ky;
        ky.add(V(bU, i + 1, jU)); ky.add(V(btU, i + 1, jU)); ky.add(V(bU, i, jU)); ky.add(V(btU, i, jU));
        ky.add(-V(bD, i + 1, jD)); ky.add(-V(btD, i + 1, jD)); ky.add(-V(bD, i, jD)); ky.add(-V(btD, i, jD));
        const double dbdy = 0.25 * p.dyr * ky.s;
        double ib = V(bt, i, j) + V(bt, i + 1, j);
        ib = ib + V(b, i, j);
        ib = ib + V(b, i + 1, j);
        ib = ib * 0.5;
        const size_t fi = (size_t)t * p.NFY + f;
        const double fg = gamma_of(p.geometric, dbdx, dbdy);
        d.yb0[fi] = 0.5 * (V(b, i, j) + V(b, i + 1, j));
        T.ybt[fi] = btf;
        T.ybx[fi] = dbdx;
        T.ygam[fi] = fg;
        T.yib[fi] = ib;
        if (STATIC) {
            const double a = fmin(gl / fg, 1.0), c = fmin(gr_ / fg, 1.0);
            T.ynp[fi] = a * a * p.dy;
            T.ynm[fi] = c * c * p.dy;
        }
    }
}
#undef V

// Per-face CFL numerators and per-cell surface-gradient prefactors of tile t.
// gam_ratio = min(gam(cell)/gam(face), 1) of the cell on each side of a face
// (src/HydraulicRHS.f90:983-1006) enters the CFL as (gr*gr)*dx; the
// prefactors of ConstructHydraulicRHS (src/HydraulicRHS.f90:1233-1243).
// Needs the cell gamma of the neighbour tiles.
template <class DevT, class TopoT>
__device__ void tile_prefactors(const DevT& d, const Params& p, const TopoT& T, int t, int part = 0,
                                int nparts = 1) {
    const int lane0 = threadIdx.x + part * blockDim.x, stride = blockDim.x * nparts;
    for (int c = lane0; c < p.NC; c += stride) {
        const int gi = t * p.NC + c;
        const double bx = T.cbx[gi], by = T.cby[gi], gam = T.cgam[gi];
        if (p.geometric) {
            T.pxu[gi] = (1.0 + by * by) / gam;
            T.pxv[gi] = -(bx * by / gam);
            T.pyv[gi] = (1.0 + bx * bx) / gam;
        } else {
            T.pxu[gi] = 1.0; T.pxv[gi] = 0.0; T.pyv[gi] = 1.0;
        }
    }
}

template <class DevT, class TopoT>
__device__ void tile_static_terms(const DevT& d, const Params& p, const TopoT& T, int t, int part = 0,
                                  int nparts = 1) {
    const int lane0 = threadIdx.x + part * blockDim.x, stride = blockDim.x * nparts;
    const int nx = p.nx, ny = p.ny;
    int tW, tE, tS, tN;
    side_tiles(p, t, tW, tE, tS, tN);
    tile_prefactors(d, p, T, t, part, nparts);
    for (int f = lane0; f < (nx + 1) * ny; f += stride) {
        const int i = f % (nx + 1), j = f / (nx + 1);
        const double fg = T.xgam[t * p.NFX + f];
        const double gl = (i == 0) ? T.cgam[tW * p.NC + j * nx + nx - 1] : T.cgam[t * p.NC + j * nx + i - 1];
        const double gr_ = (i == nx) ? T.cgam[tE * p.NC + j * nx] : T.cgam[t * p.NC + j * nx + i];
        const double a = fmin(gl / fg, 1.0), b = fmin(gr_ / fg, 1.0);
        T.xnp[t * p.NFX + f] = a * a * p.dx;
        T.xnm[t * p.NFX + f] = b * b * p.dx;
    }
    for (int f = lane0; f < (p.oned ? 0 : nx * (ny + 1)); f += stride) {
        const int i = f % nx, j = f / nx;
        const double fg = T.ygam[t * p.NFY + f];
        const double gl = (j == 0) ? T.cgam[tS * p.NC + (ny - 1) * nx + i] : T.cgam[t * p.NC + (j - 1) * nx + i];
        const double gr_ = (j == ny) ? T.cgam[tN * p.NC + i] : T.cgam[t * p.NC + j * nx + i];
        const double a = fmin(gl / fg, 1.0), b = fmin(gr_ / fg, 1.0);
        T.ynp[t * p.NFY + f] = a * a * p.dy;
        T.ynm[t * p.NFY + f] = b * b * p.dy;
    }
}

// ---------------------------------------------------------------------------
// tile activation (src/TimeStepper.f90:924-1150, src/UpdateTiles.f90:56-481)
// ---------------------------------------------------------------------------

// Ghost w slopes at the edges of tile g that face active tiles
// (ReconstructwAtEdges, src/UpdateTiles.f90:669-750).
template <class DevT>
__device__ void ghost_slopes(const DevT& d, const Params& p, int g) {
    const int nx = p.nx, ny = p.ny, NC = p.NC;
    const double* w = (d.mixed && d.ctrl->cur) ? d.S0b : d.S0;
    const size_t base = (size_t)g * NC;
    auto on = [&](int k) { return k >= 0 && k < p.ntiles && d.tstate[k] == KST_TILE_ACTIVE; };
    int nb = nb_tile(p, g, -1, 0);
    if (on(nb))
        for (int j = threadIdx.x; j < ny; j += blockDim.x) {
            const double c = w[base + j * nx], e = w[base + j * nx + 1];
            const double wv = w[(size_t)nb * NC + j * nx + nx - 1];
            d.gsx[base + j * nx] = p.dxr * limiter(p.limiter, e - c, c - wv);
        }
    nb = nb_tile(p, g, 1, 0);
    if (on(nb))
        for (int j = threadIdx.x; j < ny; j += blockDim.x) {
            const double c = w[base + j * nx + nx - 1], we = w[base + j * nx + nx - 2];
            const double ev = w[(size_t)nb * NC + j * nx];
            d.gsx[base + j * nx + nx - 1] = p.dxr * limiter(p.limiter, ev - c, c - we);
        }
    nb = nb_tile(p, g, 0, -1);
    if (on(nb))
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const double c = w[base + i], n = w[base + nx + i];
            const double sv = w[(size_t)nb * NC + (ny - 1) * nx + i];
            d.gsy[base + i] = p.dyr * limiter(p.limiter, n - c, c - sv);
        }
    nb = nb_tile(p, g, 0, 1);
    if (on(nb))
        for (int i = threadIdx.x; i < nx; i += blockDim.x) {
            const double c = w[base + (ny - 1) * nx + i], sv = w[base + (ny - 2) * nx + i];
            const double nv = w[(size_t)nb * NC + i];
            d.gsy[base + (ny - 1) * nx + i] = p.dyr * limiter(p.limiter, nv - c, c - sv);
        }
}

// The state of a freshly activated tile (ActivateTile sets w = b0 over
// AllocateU's or the ghost data's zero momenta; src/UpdateTiles.f90:342-384)
// or of a new ghost tile (UpdateGhostTileBCs, src/UpdateTiles.f90:571-585):
// on the domain edge with Dirichlet boundaries the boundary data
// (SetDomainBoundaryData, :595-643), with periodic ones none (AllocateU's
// zeros, w = 0 included), else the dry data of SetDefaultTileData (:647-664).
struct CellData {
    double w, ru, rv, hp;
};

template <class DevT>
__device__ __forceinline__ CellData new_tile_cell(const DevT& d, const Params& p, int k, size_t gi,
                                                 bool active) {
    const double b = d.b0c[gi];
    if (!active && p.bc != KST_BC_HALT && edge_tile(p, k)) {
        if (p.bc == KST_BC_PERIODIC) return {0.0, 0.0, 0.0, 0.0};
        if (p.bc == KST_BC_DIRICHLET) {
            const double hn = p.bc_hn;
            const double hp = hn / d.T.cgam[gi];
            double w = d.T.btc[gi] + hp;
            w = w + b;
            return {w, p.bc_rho * hn * p.bc_u, p.bc_rho * hn * p.bc_v, hn * p.bc_psi};
        }
    }
    return {b, 0.0, 0.0, 0.0};
}

template <class DevT>
__device__ void dry_tile(const DevT& d, const Params& p, int k, bool active) {
    const size_t base = (size_t)k * p.NC;
    const long long FS = d.FS;
    for (int c = threadIdx.x; c < p.NC; c += blockDim.x) {
        const CellData v = new_tile_cell(d, p, k, base + c, active);
        if (d.mixed) {
            // both FP64 states hold the new data; increments and RHS are zero
            double* S[2] = {d.S0, d.S0b};
            for (int q = 0; q < 2; ++q) {
                S[q][base + c] = v.w;
                S[q][FS + base + c] = v.ru;
                S[q][2 * FS + base + c] = v.rv;
                S[q][3 * FS + base + c] = v.hp;
            }
            float* S2f = (float*)d.S2;
            float* S3f = (float*)d.S3;
            float* Rf = (float*)d.R;
            for (int q = 0; q < 4; ++q) { S2f[q * FS + base + c] = 0.f; S3f[q * FS + base + c] = 0.f; }
            for (int q = 0; q < 5; ++q) Rf[q * FS + base + c] = 0.f;
            continue;
        }
        double* S[4] = {d.S0, d.S2, d.S3, d.SH};
        for (int q = 0; q < (d.morpho ? 4 : 3); ++q) {
            S[q][base + c] = v.w;
            S[q][FS + base + c] = v.ru;
            S[q][2 * FS + base + c] = v.rv;
            S[q][3 * FS + base + c] = v.hp;
        }
        for (int q = 0; q < 5; ++q) d.R[q * FS + base + c] = 0.0;
        if (d.morpho) {
            for (int q = 0; q < 5; ++q) d.R2[q * FS + base + c] = 0.0;
            for (int q = 0; q < 3; ++q) { d.SM[q][base + c] = v.w; d.SM[q][FS + base + c] = v.hp; }
            d.UV[base + c] = 0.0; d.UV[FS + base + c] = 0.0;
        }
        if (active) {
            for (int q = 0; q < 10; ++q) d.maxima[q * FS + base + c] = 0.0;
            d.maxima[10 * FS + base + c] = -1.0;
        }
    }
}

static constexpr int MAXNEW = 1024;

// A new step refreshes intermed3's ghost tiles from the tileContainer
// (InitialiseTimeSteppingArrays, src/TimeStepper.f90:785-808): undo the
// redistribution writes of the previous step.
template <class DevT>
__device__ void refresh_ghost_overlay(const DevT& d, const Params& p) {
    Ctrl* c = d.ctrl;
    const long long n = c->ndirty;
    if (n == 0) return;
    const long long FS = d.FS;
    const size_t NV = (size_t)(p.nx + 1) * (p.ny + 1);
    const long long vbit = 1LL << 62;
    if (n <= d.dirty_cap) {
        for (long long q = threadIdx.x; q < n; q += blockDim.x) {
            const long long e = d.dirty[q];
            if (e & vbit) {
                const size_t v = (size_t)(e & ~vbit);
                d.TB.vbt[v] = d.TA.vbt[v];
            } else {
                d.S3[e] = d.S0[e];
                d.S3[3 * FS + e] = d.S0[3 * FS + e];
            }
        }
    } else {
        for (int g = 0; g < c->n_ghost; ++g) {
            const int k = d.ghosts[g];
            for (size_t q = threadIdx.x; q < NV; q += blockDim.x) d.TB.vbt[k * NV + q] = d.TA.vbt[k * NV + q];
            for (int q = threadIdx.x; q < p.NC; q += blockDim.x) {
                const size_t gi = (size_t)k * p.NC + q;
                d.S3[gi] = d.S0[gi];
                d.S3[3 * FS + gi] = d.S0[3 * FS + gi];
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) c->ndirty = 0;
    __syncthreads();
}

// Neighbour r of tile k in the order of the activation passes: N, S, E, W
// (NearNorthBoundary .. NearWestBoundary, src/TimeStepper.f90:946-1150).
__device__ __forceinline__ int side_tile(const Params& p, int k, int r) {
    const int di = r == 2 ? 1 : (r == 3 ? -1 : 0), dj = r == 0 ? 1 : (r == 1 ? -1 : 0);
    return nb_tile(p, k, di, dj);
}

// side_tile(p, k, r) for r = 0..3 (N, S, E, W), with one division at most.
__device__ __forceinline__ void pass_sides(const Params& p, int k, int (&n)[4]) {
    if (!p.periodic) {
        n[0] = k + p.ntx; n[1] = k - p.ntx; n[2] = k + 1; n[3] = k - 1;
        return;
    }
    const int tj = k / p.ntx, ti = k - tj * p.ntx;
    const int row = tj * p.ntx;
    n[0] = ti + (tj + 1 >= p.nty ? tj + 1 - p.nty : tj + 1) * p.ntx;
    n[1] = ti + (tj - 1 < 0 ? tj - 1 + p.nty : tj - 1) * p.ntx;
    n[2] = row + (ti + 1 >= p.ntx ? ti + 1 - p.ntx : ti + 1);
    n[3] = row + (ti - 1 < 0 ? ti - 1 + p.ntx : ti - 1);
}

// What AddTile (src/UpdateTiles.f90:56-78) does with neighbour n of an
// active tile whose edge the flow nears: nothing if n is no tile (id -1 in
// the reference's 0-based terms) or already active; outside the domain or on
// its edge (but with periodic boundaries) a halt with 'halt' boundaries and
// nothing otherwise; else activation.
enum { ADD_NONE = 0, ADD_TILE = 1, ADD_HALT = 2 };
__device__ __forceinline__ int add_action(const Params& p, const uint8_t* ts, int n) {
    if (n <= -1 || (n < p.ntiles && ts[n] == KST_TILE_ACTIVE)) return ADD_NONE;
    if (n >= p.ntiles || (!p.periodic && edge_tile(p, n)))
        return p.bc == KST_BC_HALT ? ADD_HALT : ADD_NONE;
    return ADD_TILE;
}

// The action for neighbour n given its state byte (read by the caller when
// 0 <= n < ntiles).
__device__ __forceinline__ int add_action_of(const Params& p, int n, uint8_t tsn) {
    if (n <= -1 || (n < p.ntiles && tsn == KST_TILE_ACTIVE)) return ADD_NONE;
    if (n >= p.ntiles || (!p.periodic && edge_tile(p, n)))
        return p.bc == KST_BC_HALT ? ADD_HALT : ADD_NONE;
    return ADD_TILE;
}

// The ghost candidates of a new tile (AddGhostTiles, src/UpdateTiles.f90:389-481):
// W, E, N, S and, but with periodic boundaries, the four diagonal tiles; W
// and E on a 1D grid.
__device__ __forceinline__ int ghost_count(const Params& p) { return p.oned ? 2 : (p.periodic ? 4 : 8); }
__device__ __forceinline__ int ghost_candidate(const Params& p, int k, int r) {
    constexpr int DI[8] = {-1, 1, 0, 0, -1, 1, -1, 1}, DJ[8] = {0, 0, 1, -1, -1, -1, 1, 1};
    return nb_tile(p, k, DI[r], DJ[r]);
}

static constexpr int ACT_NT = 256;      // activation workgroup: 4 waves
static constexpr int ACT_R = 8;         // active-list entries per thread and round
static constexpr int MAXCAND = 2048;    // candidate tiles held in LDS

// Ordered append over the workgroup: each thread brings `cnt` items; the
// return value is the first slot of this thread's items, slots following
// thread order after the `base` slots already taken; `base` (uniform)
// grows by the round's total.
__device__ __forceinline__ int block_append(int cnt, int& base, int* wsum) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int off = base, tot = 0;
    for (int w = 0; w < ACT_NT / 64; ++w) {
        if (w < wv) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    base += tot;
    return off + x - cnt;
}

// The N, S, E, W passes of CheckIfNearBoundaries over the active list in
// place (src/TimeStepper.f90:935-1150): a new tile is inserted in order and
// the later entries of the pass shift, so entries pushed past the pass's
// initial length are not visited.  One thread walks the list; taken when the
// candidates do not fit in LDS.  eflags hold each tile's candidate sides.
template <class DevT>
__device__ void activate_walk(const DevT& d, const Params& p, int* newt, int& nnew, int& bad) {
    Ctrl* c = d.ctrl;
    for (int q = p.oned ? 2 : 0; q < 4 && !bad; ++q) {
        const int size = c->n_active;
        for (int tt = 0; tt < size && !bad; ++tt) {
            const int k = d.active[tt];
            const int n = side_tile(p, k, q);
            if (!(d.eflags[k] & (1u << q))) continue;
            const int act = add_action(p, d.tstate, n);
            if (act == ADD_HALT) { bad = 1; break; }
            if (act == ADD_TILE) {
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
}

// Activation after an accepted step (CheckIfNearBoundaries,
// src/TimeStepper.f90:935-1150; AddTile, src/UpdateTiles.f90:56-78).
// A flagged side of an active tile is a candidate when its neighbour would
// be activated or halt the run; sides that are not candidates at the start
// never act later (tiles only become active), so the walk over the list
// reduces to a walk over the candidates in list order: a candidate at
// position i0 of the initial list sits at i0 + (new tiles below it) when
// the pass reaches it, and is visited if that is below the pass's initial
// length.  The new tiles then merge into the sorted list, and leave the
// ghost list in order, in parallel.
template <class DevT>
__device__ __forceinline__ void activate_body(const DevT& d, const Params& p) {
    __shared__ int newt[MAXNEW], srt[MAXNEW], ipos[MAXNEW];
    __shared__ int cand_k[MAXCAND], cand_e[MAXCAND];
    __shared__ int wsum[ACT_NT / 64], tl[ACT_NT];
    __shared__ int nnew, bad, first_move, ng_start;
    Ctrl* c = d.ctrl;
    if (c->status != ST_NEW) return;
    if (threadIdx.x == 0) count_work(c, KW_ACTIVATE);
    if (d.morpho) refresh_ghost_overlay(d, p);
    const int n0 = c->n_active;
    const int s0 = p.oned ? 2 : 0;      // E and W only on a 1D grid
    // candidates in list order; each thread takes ACT_R consecutive entries
    // with every load of the round in flight at once
    int ncand = 0;
    for (int base = 0; base < n0; base += ACT_NT * ACT_R) {
        const int q0 = base + threadIdx.x * ACT_R;
        int k[ACT_R];
        unsigned f[ACT_R];
        // the loads take clamped (in-range) indices unconditionally and the
        // results are selected afterwards, so that a round's loads issue
        // back to back rather than each behind a branch and a wait
#pragma unroll
        for (int r = 0; r < ACT_R; ++r) {
            const int v = d.active[min(q0 + r, n0 - 1)];
            k[r] = q0 + r < n0 ? v : -1;
        }
#pragma unroll
        for (int r = 0; r < ACT_R; ++r) {
            const unsigned v = d.eflags[max(k[r], 0)];
            f[r] = k[r] >= 0 ? v : 0u;
        }
        int nb[ACT_R][4], ts[ACT_R][4];
#pragma unroll
        for (int r = 0; r < ACT_R; ++r) {
            pass_sides(p, max(k[r], 0), nb[r]);
#pragma unroll
            for (int s = 0; s < 4; ++s) ts[r][s] = d.tstate[min(max(nb[r][s], 0), p.ntiles - 1)];
        }
        unsigned b[ACT_R];
        int cnt = 0;
#pragma unroll
        for (int r = 0; r < ACT_R; ++r) {
            b[r] = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (s >= s0 && ((f[r] >> s) & 1u) &&
                    add_action_of(p, nb[r][s], (uint8_t)ts[r][s]) != ADD_NONE)
                    b[r] |= 1u << s;
            cnt += b[r] != 0;
        }
        int slot = block_append(cnt, ncand, wsum);
        // the flags become the candidate sides after the barriers of the
        // append, which would otherwise wait for these stores
#pragma unroll
        for (int r = 0; r < ACT_R; ++r) {
            if (f[r]) d.eflags[k[r]] = b[r];
            if (b[r]) {
                if (slot < MAXCAND) { cand_k[slot] = k[r]; cand_e[slot] = ((q0 + r) << 4) | (int)b[r]; }
                ++slot;
            }
        }
    }
    if (ncand == 0) return;   // every flag cleared, nothing to add
    __syncthreads();
    if (ncand > d.act_cand) {
        if (threadIdx.x == 0) {
            nnew = 0;
            bad = 0;
            activate_walk(d, p, newt, nnew, bad);
            if (bad) {
                c->status = ST_ERROR;
                c->err = (bad == 1) ? KST_ERR_HALT : KST_ERR_STATE;
            }
        }
        __syncthreads();
        if (bad) return;
        for (int q = threadIdx.x; q < c->n_active; q += ACT_NT) d.eflags[d.active[q]] = 0;
    } else {
        if (threadIdx.x == 0) {
            nnew = 0;
            bad = 0;
            int size = n0;
            for (int s = s0; s < 4 && !bad; ++s) {
                for (int j = 0; j < ncand; ++j) {
                    const int e = cand_e[j];
                    if (!(e & (1 << s))) continue;
                    const int v = cand_k[j];
                    int below = 0;
                    for (int m = 0; m < nnew; ++m) below += newt[m] < v;
                    if ((e >> 4) + below >= size) break;
                    const int n = side_tile(p, v, s);
                    const int act = add_action(p, d.tstate, n);
                    if (act == ADD_HALT) { bad = 1; break; }
                    if (act == ADD_TILE) {
                        d.tstate[n] = KST_TILE_ACTIVE;
                        d.eflags[n] = 0;
                        if (nnew == MAXNEW) { bad = 2; break; }
                        newt[nnew++] = n;
                    }
                }
                size = n0 + nnew;
            }
            if (bad) {
                c->status = ST_ERROR;
                c->err = (bad == 1) ? KST_ERR_HALT : KST_ERR_STATE;
            }
        }
        __syncthreads();
        if (bad) return;
        for (int j = threadIdx.x; j < ncand; j += ACT_NT) d.eflags[cand_k[j]] = 0;
        const int nn = nnew;
        if (nn == 0) return;
        // the new tiles in order (ranks by counting: the values are distinct)
        // and each one's place in the initial list
        for (int m = threadIdx.x; m < nn; m += ACT_NT) {
            const int v = newt[m];
            int rk = 0;
            for (int i = 0; i < nn; ++i) rk += newt[i] < v;
            srt[rk] = v;
            int lo = 0, hi = n0;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (d.active[mid] < v) lo = mid + 1;
                else hi = mid;
            }
            ipos[rk] = lo;
            if (rk == 0) first_move = lo;
        }
        __syncthreads();
        // move the initial entries up by the new tiles below them, the top
        // round first (every entry is read before any is written)
        for (int top = ((n0 - 1) / (ACT_NT * ACT_R)) * (ACT_NT * ACT_R); top + ACT_NT * ACT_R > first_move;
             top -= ACT_NT * ACT_R) {
            int v[ACT_R];
#pragma unroll
            for (int r = 0; r < ACT_R; ++r) {
                const int q = top + r * ACT_NT + (int)threadIdx.x;
                v[r] = q < n0 && q >= first_move ? d.active[q] : -1;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < ACT_R; ++r)
                if (v[r] >= 0) {
                    int lo = 0, hi = nn;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (srt[mid] < v[r]) lo = mid + 1;
                        else hi = mid;
                    }
                    d.active[top + r * ACT_NT + threadIdx.x + lo] = v[r];
                }
            __syncthreads();
            if (top == 0) break;
        }
        for (int m = threadIdx.x; m < nn; m += ACT_NT) d.active[ipos[m] + m] = srt[m];
        // tiles that were ghosts leave the ghost list, which keeps its order
        const int ng0 = c->n_ghost;
        int kept = 0;
        for (int base = 0; base < ng0; base += ACT_NT) {
            const int q = base + threadIdx.x;
            const int g = q < ng0 ? d.ghosts[q] : -1;
            const bool keep = g >= 0 && d.tstate[g] != KST_TILE_ACTIVE;
            const int slot = block_append(keep ? 1 : 0, kept, wsum);
            if (keep) d.ghosts[slot] = g;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            c->n_active = n0 + nn;
            c->n_ghost = kept;
        }
        __syncthreads();
    }
    if (nnew == 0) return;
    __syncthreads();
    // ActivateTile (src/UpdateTiles.f90:342-384): cell bt = 0, face
    // topography from the tile's and its neighbours' current vertex bt
    for (int q = 0; q < nnew; ++q) {
        tile_cell_topo(d, p, d.T, newt[q], true);
        __syncthreads();
        tile_face_topo(d, p, d.T, newt[q]);
        __syncthreads();
        tile_static_terms(d, p, d.T, newt[q]);
        __syncthreads();
    }
    for (int q = 0; q < nnew; ++q) dry_tile(d, p, newt[q], true);
    // new ghosts around each new tile (AddGhostTiles, src/UpdateTiles.f90:389-481)
    if (threadIdx.x == 0) {
        ng_start = c->n_ghost;
        for (int q = 0; q < nnew; ++q) {
            const int k = newt[q];
            for (int r = 0; r < ghost_count(p); ++r) {
                const int g = ghost_candidate(p, k, r);
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
    // ghosts whose neighbourhood changed rebuild their edge w slopes, taken
    // in list order a round of ACT_NT at a time
    const int ngh = c->n_ghost;
    for (int base = 0; base < ngh; base += ACT_NT) {
        const int q = base + threadIdx.x;
        const int g = q < ngh ? d.ghosts[q] : -1;
        bool touched = q < ngh && q >= ng_start;
        for (int r = 0; r < nnew && g >= 0 && !touched; ++r)
            for (int u = 0; u < ghost_count(p) && !touched; ++u)
                if (ghost_candidate(p, newt[r], u) == g) touched = true;
        int cnt = 0;
        const int slot = block_append(touched ? 1 : 0, cnt, wsum);
        if (touched) tl[slot] = g;
        __syncthreads();
        for (int j = 0; j < cnt; ++j) ghost_slopes(d, p, tl[j]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(ACT_NT) void activate_kernel(const Dev* __restrict__ d,
                                                          const Params* __restrict__ p) {
    activate_body(const_arg(d), *p);
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
        const double* w = (d.mixed && d.ctrl->cur) ? d.S0b : d.S0;
        double hn = compute_hn(w[gi], d.b0c[gi], d.T.btc[gi], d.T.cgam[gi]);
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

// Accept or roll back the step (src/TimeStepper.f90:173-275).  With
// morphodynamics the attempt is H(dt) M(2dt) H(dt) and fails at the first
// of: a CFL re-check of the first operator (dt <- 0.9 dt_k), the
// morphodynamic checks or redistribution (dt <- dt/2), the second
// operator's substep-1 advice (dt <- 0.9 dt or the advice), or its CFL
// re-checks (dt <- 0.9 dt_k).
template <class DevT>
__device__ void finalize_body(const DevT& d, const Params& p) {
    Ctrl* c = d.ctrl;
    c->commit = 0;
    if (c->status == ST_DONE || c->status == ST_ERROR) return;
    c->work[KW_FINALIZE] += 1;
    const bool ran_p1 = (c->status == ST_NEW);
    const double dt = step_dt(d, p, d.ctrl);
    c->rhs_evals += (ran_p1 ? 1 : 0) + 1;
    bool reject = false;
    double retry = 0.0;
    const double dt1 = advised_from_slot(p, c->slot[1].bits, c->n_active);
    if (dt1 < dt) {
        reject = true;
        retry = 0.9 * dt1;
    } else {
        c->rhs_evals += 1;
        const double dt2 = advised_from_slot(p, c->slot[2].bits, c->n_active);
        if (dt2 < dt) {
            reject = true;
            retry = 0.9 * dt2;
        } else {
            c->rhs_evals += 1;
        }
    }
    if (d.morpho && !reject) {
        if (c->mref) {
            reject = true;
            retry = 0.5 * dt;
        } else {
            c->rhs_evals += 1;
            const double adv = 0.9 * advised_from_slot(p, c->slot[3].bits, c->n_active);
            if (adv < dt) {
                reject = true;
                retry = ((dt - adv) / dt < 0.1) ? 0.9 * dt : adv;
            } else {
                c->rhs_evals += 1;
                const double e1 = advised_from_slot(p, c->slot[4].bits, c->n_active);
                if (e1 < dt) {
                    reject = true;
                    retry = 0.9 * e1;
                } else {
                    c->rhs_evals += 1;
                    const double e2 = advised_from_slot(p, c->slot[5].bits, c->n_active);
                    if (e2 < dt) {
                        reject = true;
                        retry = 0.9 * e2;
                    } else {
                        c->rhs_evals += 1;
                    }
                }
            }
        }
    }
    const unsigned long long huge = dbits(kHuge);
    c->mref = 0;
    c->ncand = 0;
    for (int k = 1; k < 6; ++k) c->slot[k].bits = huge;
    if (reject) {
        c->rejected++;
        c->status = ST_RETRY;
        c->dt_retry = retry;
        return;
    }
    c->t = d.morpho ? c->t + 2.0 * dt : c->t + dt;
    c->next_t = next_flux_time(d, p, c->t);
    c->steps++;
    c->cell_updates += (long long)c->n_active * p.NC;
    c->last_dt = dt;
    if (c->steps == 1 || dt < c->min_dt) c->min_dt = dt;
    if (dt > c->max_dt) c->max_dt = dt;
    c->slot[0].bits = huge;
    c->status = ST_NEW;
    c->commit = d.morpho ? 1 : 0;
    if (d.mixed) c->cur ^= 1;
    if (c->mode == 0) {
        if (c->t >= c->tend) c->status = ST_DONE;
    } else {
        if (--c->steps_left <= 0) c->status = ST_DONE;
    }
}

__global__ void finalize_kernel(const Dev* __restrict__ d, const Params* __restrict__ p) {
    finalize_body(const_arg(d), *p);
}

// ---------------------------------------------------------------------------
// batches: many scenarios advanced by one launch per kernel
// ---------------------------------------------------------------------------

// A slot table of resident scenarios (each with its own state, tiles, dt and
// control block) and the work list of one iteration: off[s] is the first
// work item (sub-block of an active tile) of slot s, off[nslot] the total.
// Tiles' per-tile passes over the bed (face topography of a new bed, the
// accepted bed's commit) are split into this many (tile, part) items.
constexpr int TILE_PARTS = 8;

struct Batch {
    const Dev* dev;
    const Params* par;
    const int* on;        // slot holds a scenario
    int* off;             // [nslot + 1] first work item of each slot (16 x 16 sub-blocks)
    int* status;          // [nslot] the scenario's status after each iteration
    int nslot;
    const Dev* dev2;      // the second hydraulic operator of each slot (morphodynamic slots)
    int* off2;            // [nslot + 1] as off, morphodynamic slots only
    int* off3;            // [nslot + 1] (tile, part) items of the morphodynamic slots
};

constexpr int BATCH_PLAN_THREADS = 1024;

// Activation of every slot (one block per slot).
__global__ __launch_bounds__(ACT_NT) void batch_activate_kernel(Batch bt) {
    const int s = blockIdx.x;
    if (!bt.on[s]) return;
    activate_body(const_arg(bt.dev + s), generic(const_arg(bt.par + s)));
}

// Work list of the iteration: the running slots' active sub-blocks.
// The iteration's work lists (exclusive prefix sums over the slots): off,
// the running slots' sub-blocks; with morphodynamic slots (lists > 1) also
// off2, theirs alone, and off3, their (tile, part) items.
__global__ __launch_bounds__(BATCH_PLAN_THREADS) void batch_plan_kernel(Batch bt, int lists) {
    __shared__ int sc[BATCH_PLAN_THREADS];
    for (int kind = 0; kind < lists; ++kind) {
        int* off = kind == 0 ? bt.off : (kind == 1 ? bt.off2 : bt.off3);
        int carry = 0;
        for (int base = 0; base < bt.nslot; base += BATCH_PLAN_THREADS) {
            const int s = base + (int)threadIdx.x;
            int n = 0;
            if (s < bt.nslot && bt.on[s]) {
                const Dev& d = bt.dev[s];
                const Params& p = bt.par[s];
                const Ctrl* c = d.ctrl;
                const bool run = c->status == ST_NEW || c->status == ST_RETRY;
                const int spt = ((p.nx + B - 1) / B) * ((p.ny + B - 1) / B);
                if (kind == 0 && run) n = c->n_active * spt;
                if (kind == 1 && run && d.morpho) n = c->n_active * spt;
                if (kind == 2 && d.morpho) n = c->n_active * TILE_PARTS;
            }
            sc[threadIdx.x] = n;
            __syncthreads();
            for (int o = 1; o < BATCH_PLAN_THREADS; o <<= 1) {
                const int v = threadIdx.x >= (unsigned)o ? sc[threadIdx.x - o] : 0;
                __syncthreads();
                sc[threadIdx.x] += v;
                __syncthreads();
            }
            if (s < bt.nslot) off[s] = carry + sc[threadIdx.x] - n;
            carry += sc[BATCH_PLAN_THREADS - 1];
            __syncthreads();
        }
        if (threadIdx.x == 0) off[bt.nslot] = carry;
        __syncthreads();
    }
}

// One RK stage over the work list.  The grid is a multiple of 8 blocks that
// stay resident and walk the list; XCD x (blocks x, x+8, ...) takes the x-th
// contiguous eighth of it, so neighbouring sub-blocks share an L2.
// second: the second hydraulic operator of the morphodynamic slots (dev2, off2)
template <int STAGE, bool VISC, int LIM, bool MIX>
__global__ __launch_bounds__(STAGE_NT) void batch_stage_kernel(Batch bt, int second) {
    __shared__ SmemT<VISC ? 7 : 5, MIX> s;
    __shared__ int sh_slot;
    const int* off = second ? bt.off2 : bt.off;
    const Dev* dev = second ? bt.dev2 : bt.dev;
    const int total = off[bt.nslot];
    const int m = gridDim.x / 8, x = blockIdx.x % 8;
    const int per = (total + 7) / 8;
    const int end = min((x + 1) * per, total);
    for (int w = x * per + (int)(blockIdx.x / 8); w < end; w += m) {
        if (threadIdx.x == 0) {
            int lo = 0, hi = bt.nslot - 1;       // last slot with off <= w
            while (lo < hi) {
                const int mid = (lo + hi + 1) / 2;
                if (off[mid] <= w) lo = mid; else hi = mid - 1;
            }
            sh_slot = lo;
        }
        __syncthreads();
        const int sl = __builtin_amdgcn_readfirstlane(sh_slot);
        stage_body<STAGE, VISC, LIM, MIX>(const_arg(dev + sl), generic(const_arg(bt.par + sl)), s,
                                          w - off[sl]);
        __syncthreads();
    }
}

// Step control of every slot, and the status table the host polls.
__global__ void batch_finalize_kernel(Batch bt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= bt.nslot || !bt.on[s]) return;
    finalize_body(bt.dev[s], bt.par[s]);
    bt.status[s] = bt.dev[s].ctrl->status;
}

// Topography of every tile of a set (after kst_set_topography / kst_set_state).
__global__ __launch_bounds__(256) void topo_kernel(Dev d, Params p, Topo T) {
    tile_cell_topo(d, p, T, blockIdx.x, false);
    __syncthreads();
    tile_face_topo(d, p, T, blockIdx.x);
}

// Needs the cell gamma of the neighbour tiles, so it runs after topo_kernel.
__global__ __launch_bounds__(256) void topo2_kernel(Dev d, Params p, Topo T) {
    tile_static_terms(d, p, T, blockIdx.x);
}

// Cell gamma of tile blockIdx.x from its cell slopes (after
// kst_set_cell_topography replaced some of them); the form of tile_cell_topo.
__global__ __launch_bounds__(256) void cell_gamma_kernel(Dev d, Params p, Topo T) {
    const size_t base = (size_t)blockIdx.x * p.NC;
    for (int c = threadIdx.x; c < p.NC; c += blockDim.x)
        T.cgam[base + c] = gamma_of(p.geometric, T.cbx[base + c], p.oned ? 0.0 : T.cby[base + c]);
}

}  // namespace kst

#include "kst_morpho.hpp"
#include "kst_output.hpp"

// ===========================================================================
// host side: the C-ABI (include/kestrel_step.h)
// ===========================================================================
using namespace kst;

struct kst_handle {
    kst_config cfg;
    Params p;
    Dev d;          // launch arguments of the first (or only) hydraulic operator
    Dev d2;         // of the second hydraulic operator of a Strang step
    Dev* d_args = nullptr;      // device copies of d and d2
    Params* d_par = nullptr;    // device copy of p
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    std::string err;
    int nx, ny, ntx, nty, ntiles, NX, NY, NC;
    bool have_topo = false, have_state = false;
    std::vector<std::vector<double>> src_times;   // flux-source knot times (host copy)
    double t = 0.0;
    // profiling: per-iteration event pairs around the four stage kernels
    bool profiling = false;
    // kGraphIters whole iterations, per (grid cap, sub-block shape)
    std::vector<std::pair<long long, hipGraphExec_t>> graphs;
    int fixed_cap = 0;                   // kst_set_grid_cap (0: from the active set)
    int resident = 0;                    // 16 x 16 stage-kernel blocks resident on the device at once
    bool autotune = false;               // kst_autotune: choices per active-set bucket
    struct Tune {
        int small;                       // 16 x 4 sub-blocks
        int grid;                        // launch-grid limit of the 16 x 16 shape (0: the cap's grid)
        double us[4];                    // stage-1 launch time of each candidate (0: not tried)
    };
    std::map<int, Tune> tuned;           // bucket (active tiles, rounded up to a power of two)
    double* vol_part = nullptr;          // volume partials (VOL_BLOCKS_MAX + 1 groups)
    double* out_buf = nullptr;           // kst_get_output staging, out_cap doubles
    size_t out_cap = 0;
    std::vector<hipEvent_t> ev;          // 8 per iteration of the current batch
    double prof_ms[KW_N] = {};
    long long prof_n[KW_N] = {};
};

// Device scratch of kst_volume_noalloc: the partial sums of the volume
// kernels (VOL_BLOCKS_MAX groups) and their merged result (one group), the
// same layout as the handle's vol_part.  One copy per GPU, held until the
// process ends; every call on that GPU shares it.  In host code the array's
// name stands for no device address: its address on each device comes from
// hipGetSymbolAddress (scratch_ptr), taken once per device in kst_create.
constexpr size_t kScratchDoubles = (size_t)(VOL_BLOCKS_MAX + 1) * VOL_Q * 2;
static __device__ __align__(256) double scratch[kScratchDoubles];

// A batch: a fixed table of slots, each holding a scenario handle whose
// device state stays resident; one hipGraph of kBatchIters iterations walks
// every running slot (kernel arguments point at the table, so scenarios join
// and leave between launches without a new graph).
struct kst_batch {
    int nslot = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<kst_handle*> slots;
    std::vector<int> on;
    Batch bt{};
    Dev* d_dev = nullptr;
    Dev* d_dev2 = nullptr;         // each slot's second hydraulic operator
    Params* d_par = nullptr;
    int *d_on = nullptr, *d_off = nullptr, *d_status = nullptr;
    int *d_off2 = nullptr, *d_off3 = nullptr;
    int visc = -1, mixed = -1;     // fixed by the first scenario attached
    int anylim = 0;                // a running scenario's limiter is not MinMod2
    int morpho = 0;                // a running scenario has morphodynamics
    int grid = 0;                  // resident grids of the stage, morphodynamic and (tile, part) kernels
    int mgrid = 0, tgrid = 0;
    hipGraphExec_t gexec = nullptr;
    std::string err;
};

namespace {

int set_err(kst_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

// scratch's address on each device (nullptr until taken).
std::mutex g_scratch_mutex;
std::vector<double*> g_scratch_ptr;

// Take scratch's address on `dev`, which must be the current device; once
// per device.  The address is kept for every later call on that device.
int scratch_setup(kst_handle* h, int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (g_scratch_ptr[dev]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_err(h, KST_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
    return 0;
}

double* scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    return dev < (int)g_scratch_ptr.size() ? g_scratch_ptr[dev] : nullptr;
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

int alloc_topo(kst_handle* h, Topo& T) {
    const size_t FS = (size_t)h->ntiles * h->NC;
    const size_t NV = (size_t)(h->nx + 1) * (h->ny + 1) * h->ntiles;
    const size_t FX = (size_t)h->ntiles * h->p.NFX, FY = (size_t)h->ntiles * h->p.NFY;
    int rc = 0;
    rc |= dalloc(h, &T.vbt, NV);
    rc |= dalloc(h, &T.btc, FS); rc |= dalloc(h, &T.cbx, FS); rc |= dalloc(h, &T.cby, FS);
    rc |= dalloc(h, &T.cgam, FS); rc |= dalloc(h, &T.pxu, FS); rc |= dalloc(h, &T.pxv, FS);
    rc |= dalloc(h, &T.pyv, FS);
    rc |= dalloc(h, &T.xbt, FX); rc |= dalloc(h, &T.xby, FX); rc |= dalloc(h, &T.xgam, FX);
    rc |= dalloc(h, &T.xib, FX); rc |= dalloc(h, &T.xnp, FX); rc |= dalloc(h, &T.xnm, FX);
    rc |= dalloc(h, &T.ybt, FY); rc |= dalloc(h, &T.ybx, FY); rc |= dalloc(h, &T.ygam, FY);
    rc |= dalloc(h, &T.yib, FY); rc |= dalloc(h, &T.ynp, FY); rc |= dalloc(h, &T.ynm, FY);
    return rc;
}

// Topography of every tile in every set the handle holds, from the vertex
// b0 and each set's vertex bt.
int refresh_topography(kst_handle* h) {
    Topo* sets[4] = {&h->d.TA, &h->d.TB, &h->d.TM[0], &h->d.TM[1]};
    const int nsets = h->d.morpho ? 4 : 1;
    for (int q = 0; q < nsets; ++q)
        hipLaunchKernelGGL(topo_kernel, dim3(h->ntiles), dim3(256), 0, h->stream, h->d, h->p, *sets[q]);
    for (int q = 0; q < nsets; ++q)
        hipLaunchKernelGGL(topo2_kernel, dim3(h->ntiles), dim3(256), 0, h->stream, h->d, h->p, *sets[q]);
    HIPCHK(h, hipGetLastError());
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
    // pi is 4 atan(1) (src/SetPrecision.f90:38)
    const double pi = 4.0 * std::atan(1.0);
    const double t1 = std::tan(pi / 180.0);
    p.gran_stat = (p.pmin + t1) / (1.0 - p.pmin * t1);
    p.bc = c.bc;
    p.periodic = c.bc == KST_BC_PERIODIC;
    p.bc_hn = c.bc_hn; p.bc_u = c.bc_u; p.bc_v = c.bc_v; p.bc_psi = c.bc_psi;
    p.bc_rho = c.rhow + (c.rhos - c.rhow) * c.bc_psi;   // Density (src/Closures.f90)
    p.oned = c.ny_tiles * c.ny_per_tile == 1;
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

typedef void (*StageFn)(const Dev*, const Params*);
struct StageSet { StageFn fn[4]; };

template <bool VISC, int LIM, bool MIX, int BY, bool ONED>
StageSet make_set() {
    return StageSet{{stage_kernel<1, VISC, LIM, MIX, BY, ONED>, stage_kernel<2, VISC, LIM, MIX, BY, ONED>,
                     stage_kernel<3, VISC, LIM, MIX, BY, ONED>, stage_kernel<4, VISC, LIM, MIX, BY, ONED>}};
}

template <bool MIX, int BY, bool ONED = false>
const StageSet& stage_set_of(bool visc, int limiter) {
    static const StageSet sets[2][5] = {
        {make_set<false, 0, MIX, BY, ONED>(), make_set<false, 1, MIX, BY, ONED>(),
         make_set<false, 2, MIX, BY, ONED>(), make_set<false, 3, MIX, BY, ONED>(),
         make_set<false, 4, MIX, BY, ONED>()},
        {make_set<true, 0, MIX, BY, ONED>(), make_set<true, 1, MIX, BY, ONED>(),
         make_set<true, 2, MIX, BY, ONED>(), make_set<true, 3, MIX, BY, ONED>(),
         make_set<true, 4, MIX, BY, ONED>()}};
    return sets[visc ? 1 : 0][(limiter >= 0 && limiter < 5) ? limiter : 1];
}

// Sub-blocks of 16 x 4 cells when the active set is small: four times the
// workgroups, each with one face per thread and its cells and halo loaded in
// one round, so a stage that cannot fill the chip has a shorter critical
// path (16 x 8 measured between the two); the morphodynamic kernels use the
// same shape then.  FP64 only.
constexpr int BY_SMALL = B / 4;
constexpr int SMALL_WORK = 128;       // sub-blocks of 16 x 16 below which the small shape is used

// A 1D grid always takes the 16 x 4 shape (16 cells of its one row per
// workgroup) and FP64.
const StageSet& stage_set(bool visc, int limiter, bool mixed, bool small, bool oned) {
    if (oned) return stage_set_of<false, BY_SMALL, true>(visc, limiter);
    if (mixed) return stage_set_of<true, B>(visc, limiter);
    return small ? stage_set_of<false, BY_SMALL>(visc, limiter) : stage_set_of<false, B>(visc, limiter);
}

// Launch grids cover `cap` active tiles (the kernels take further items in
// turn when the active set has outgrown it).
int launch_iteration(kst_handle* h, bool profiled, bool small, int cap, int grid_limit = 0) {
    const Params& p = h->p;
    const Dev& d = h->d;
    // while profiling, the dispatch of each launch carries the event pair of
    // its kernel id, which time-stamps the kernel's own start and end
    hipEvent_t* E = profiled ? h->ev.data() : nullptr;
#define LAUNCH(id, kern, grid, block, ...)                                                  \
    do {                                                                                    \
        if (E)                                                                              \
            hipExtLaunchKernelGGL(kern, grid, block, 0, h->stream, E[2 * (id)],             \
                                  E[2 * (id) + 1], 0, __VA_ARGS__);                          \
        else                                                                                \
            hipLaunchKernelGGL(kern, grid, block, 0, h->stream, __VA_ARGS__);                \
    } while (0)
    const int spt = ((p.nx + B - 1) / B) * ((p.ny + B - 1) / B);
    // multiples of 8 blocks, as the XCD-chunked mapping assumes
    const int nblk = 8 * ((cap * spt + 7) / 8);
    const bool visc = p.nu != 0.0;
    const bool sm = (small && !d.mixed) || p.oned;
    const int spt_s = ((p.nx + B - 1) / B) * ((p.ny + BY_SMALL - 1) / BY_SMALL);
    // a grid limit (a multiple of 8) keeps that many blocks, each taking
    // further items in turn
    const int nblk_stage = sm ? 8 * ((cap * spt_s + 7) / 8) : (grid_limit > 0 ? std::min(nblk, grid_limit) : nblk);
    const Dev* const A = h->d_args;
    const Params* const P = h->d_par;
    LAUNCH(KW_ACTIVATE, activate_kernel, dim3(1), dim3(ACT_NT), A, P);
    const StageSet& k = stage_set(visc, p.limiter, d.mixed != 0, sm, p.oned != 0);
    for (int q = 0; q < 4; ++q)
        LAUNCH(KW_STAGE + q, k.fn[q], dim3(nblk_stage), dim3(STAGE_NT), A, P);
    if (d.morpho) {
        // M(2dt): three SSP stages, redistribution, face topography of the new bed
        if (p.oned) {
            constexpr int NT = morpho_nt<BY_SMALL>();
            LAUNCH(KW_MORPHO + 0, (morpho_kernel<1, BY_SMALL, true>), dim3(nblk_stage), dim3(NT), A, P);
            LAUNCH(KW_MORPHO + 1, (morpho_kernel<2, BY_SMALL, true>), dim3(nblk_stage), dim3(NT), A, P);
            LAUNCH(KW_MORPHO + 2, (morpho_kernel<3, BY_SMALL, true>), dim3(nblk_stage), dim3(NT), A, P);
        } else if (sm) {
            constexpr int NT = morpho_nt<BY_SMALL>();
            LAUNCH(KW_MORPHO + 0, (morpho_kernel<1, BY_SMALL>), dim3(nblk_stage), dim3(NT), A, P);
            LAUNCH(KW_MORPHO + 1, (morpho_kernel<2, BY_SMALL>), dim3(nblk_stage), dim3(NT), A, P);
            LAUNCH(KW_MORPHO + 2, (morpho_kernel<3, BY_SMALL>), dim3(nblk_stage), dim3(NT), A, P);
        } else {
            constexpr int NT = morpho_nt<B>();
            LAUNCH(KW_MORPHO + 0, morpho_kernel<1>, dim3(nblk), dim3(NT), A, P);
            LAUNCH(KW_MORPHO + 1, morpho_kernel<2>, dim3(nblk), dim3(NT), A, P);
            LAUNCH(KW_MORPHO + 2, morpho_kernel<3>, dim3(nblk), dim3(NT), A, P);
        }
        LAUNCH(KW_REDIST, redistribute_kernel, dim3(1), dim3(RD_THREADS), A, P);
        LAUNCH(KW_FACETOPO, face_topo_kernel, dim3(cap * TILE_PARTS), dim3(256), A, P);
        // second hydraulic operator
        for (int q = 0; q < 4; ++q)
            LAUNCH(KW_STAGE + 4 + q, k.fn[q], dim3(nblk_stage), dim3(STAGE_NT), A + 1, P);
    }
    LAUNCH(KW_FINALIZE, finalize_kernel, dim3(1), dim3(1), A, P);
    if (d.morpho) LAUNCH(KW_COMMIT, commit_kernel, dim3(cap * TILE_PARTS), dim3(256), A, P);
#undef LAUNCH
    HIPCHK(h, hipGetLastError());
    return 0;
}

constexpr int kGraphIters = 8;

// The grid cap for an active set of n tiles: a power of two with room for
// the set to grow before the next control read-back (beyond it blocks take
// further items in turn), at least 8, at most every tile.  Room to double a
// small set (the small shape's blocks take their item from the block index
// only while the grid covers the set); a quarter's room from 256 tiles, since
// each block beyond the work only reads the control block and exits: on the
// plain 2048^2 deck (1596 of 4096 tiles active) room to double launched every
// tile, 61% of the blocks idle, and cost 2.8% of the step (measured), while
// a quarter's room cost C2's small sets 1.3%.
int work_cap(const kst_handle* h, int n) {
    if (h->fixed_cap > 0) return std::min(h->fixed_cap, h->ntiles);
    int cap = 8;
    while ((n >= 256 ? 4 * cap < 5 * n : cap < 2 * n) && cap < h->ntiles) cap <<= 1;
    return std::min(cap, h->ntiles);
}

// One hipGraph holding kGraphIters whole iterations (activation, four stage
// kernels, finalize), per grid cap and sub-block shape; kernel arguments
// never change for a handle.
int ensure_graph(kst_handle* h, bool small, int cap, int grid_limit, hipGraphExec_t* out) {
    const long long key = ((long long)cap << 32) | ((long long)grid_limit << 1) | (small ? 1 : 0);
    for (const auto& kv : h->graphs)
        if (kv.first == key) { *out = kv.second; return 0; }
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int q = 0; q < kGraphIters; ++q) {
        int rc = launch_iteration(h, false, small, cap, grid_limit);
        if (rc) {
            hipGraph_t dummy = nullptr;
            (void)hipStreamEndCapture(h->stream, &dummy);
            if (dummy) (void)hipGraphDestroy(dummy);
            return rc;
        }
    }
    HIPCHK(h, hipStreamEndCapture(h->stream, &g));
    hipGraphExec_t gx = nullptr;
    const hipError_t e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(h, e);
    h->graphs.emplace_back(key, gx);
    *out = gx;
    return 0;
}


// Autotuner (kst_autotune).  For the active set of the coming launches it
// times the first RK stage kernel -- a launch that reads u^n and writes only
// the RHS buffer, the CFL minimum and the edge flags, all of which the
// iteration's own stage 1 rewrites from the same u^n -- in each candidate
// launch shape: 16 x 16 sub-blocks on the cap's grid, the same capped at the
// resident blocks of the device (each block taking further items in turn),
// at twice that, and (FP64) 16 x 4 sub-blocks.  The control block is saved
// before the trials and restored after them, so the run is unchanged.  The
// fastest shape is kept per bucket of active-set sizes (powers of two).
int tune_bucket(int n) {
    int b = 1;
    while (b < n) b <<= 1;
    return b;
}

// Choices are cached for the process, per grid layout, operator and bucket,
// so that a fresh handle of the same configuration launches tuned at once.
std::mutex g_tune_mutex;
std::map<std::string, kst_handle::Tune> g_tune_cache;

std::string tune_key(const kst_handle* h, int bucket) {
    const Params& p = h->p;
    char k[160];
    std::snprintf(k, sizeof k, "%d %d %d %d %d %d %d %d %d", p.nx, p.ny, p.ntx, p.nty, p.nu != 0.0,
                  p.limiter, h->d.mixed, h->d.morpho, bucket);
    return k;
}

int autotune_bucket(kst_handle* h, const Ctrl& c0, int cap, kst_handle::Tune* out) {
    constexpr int kTrials = 6;
    Ctrl c = c0;
    c.status = ST_NEW;          // stage 1 of the first operator does its work
    int rc = write_ctrl(h, &c);
    if (rc) return rc;
    const Params& p = h->p;
    const bool visc = p.nu != 0.0, mixed = h->d.mixed != 0;
    const int spt = ((p.nx + B - 1) / B) * ((p.ny + B - 1) / B);
    const int nblk = 8 * ((cap * spt + 7) / 8);
    struct Cand { int small, grid; } cand[4] = {{0, 0}, {0, h->resident}, {0, 2 * h->resident}, {1, 0}};
    const int ncand = mixed ? 3 : 4;
    kst_handle::Tune t{};
    t.small = 0; t.grid = 0;
    double best = 1e300;
    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    for (int q = 0; q < ncand; ++q) {
        const Cand& k = cand[q];
        if (!k.small && k.grid > 0 && k.grid >= nblk) continue;   // no different from the cap's grid
        const StageSet& set = stage_set(visc, p.limiter, mixed, k.small != 0, false);
        const int spt_s = ((p.nx + B - 1) / B) * ((p.ny + BY_SMALL - 1) / BY_SMALL);
        const int grid = k.small ? 8 * ((cap * spt_s + 7) / 8) : (k.grid > 0 ? std::min(nblk, k.grid) : nblk);
        hipLaunchKernelGGL(set.fn[0], dim3(grid), dim3(STAGE_NT), 0, h->stream, h->d_args, h->d_par);
        HIPCHK(h, hipEventRecord(e0, h->stream));
        for (int r = 0; r < kTrials; ++r)
            hipLaunchKernelGGL(set.fn[0], dim3(grid), dim3(STAGE_NT), 0, h->stream, h->d_args, h->d_par);
        HIPCHK(h, hipEventRecord(e1, h->stream));
        HIPCHK(h, hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
        const double us = 1e3 * ms / kTrials;
        t.us[q] = us;
        if (us < best) { best = us; t.small = k.small; t.grid = k.small ? 0 : k.grid; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(h, hipGetLastError());
    rc = write_ctrl(h, &c0);
    if (rc) return rc;
    *out = t;
    return 0;
}

int run_until_done(kst_handle* h, kst_stats* st, long long max_iters) {
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    // profiling runs one iteration per control read-back, so that the
    // per-kernel work counters name the launches of the iteration that did work
    int batch = h->profiling ? 1 : kGraphIters;
    long long launched = 0;
    if (h->profiling && (int)h->ev.size() < 2 * KW_N) {
        const size_t old = h->ev.size();
        h->ev.resize(2 * KW_N);
        for (size_t q = old; q < h->ev.size(); ++q) HIPCHK(h, hipEventCreate(&h->ev[q]));
    }
    for (;;) {
        if (launched > max_iters)
            return set_err(h, KST_ERR_TOO_MANY_STEPS, "step limit exceeded before reaching the target");
        launched += batch;
        long long work_before[KW_N];
        std::memcpy(work_before, c.work, sizeof work_before);
        // the sub-block shape and grid cap for this batch, from the last
        // known active set
        const int spt = ((h->p.nx + B - 1) / B) * ((h->p.ny + B - 1) / B);
        bool small = (long long)c.n_active * spt < SMALL_WORK;
        const int cap = work_cap(h, c.n_active);
        int grid_limit = 0;
        if (h->autotune && !h->p.oned) {   // a 1D grid has one launch shape
            const int bucket = tune_bucket(c.n_active);
            auto it = h->tuned.find(bucket);
            if (it == h->tuned.end()) {
                kst_handle::Tune tu{};
                const std::string key = tune_key(h, bucket);
                bool cached = false;
                {
                    std::lock_guard<std::mutex> lk(g_tune_mutex);
                    auto ct = g_tune_cache.find(key);
                    if (ct != g_tune_cache.end()) { tu = ct->second; cached = true; }
                }
                if (!cached) {
                    rc = autotune_bucket(h, c, cap, &tu);
                    if (rc) return rc;
                    std::lock_guard<std::mutex> lk(g_tune_mutex);
                    g_tune_cache[key] = tu;
                }
                it = h->tuned.emplace(bucket, tu).first;
            }
            small = it->second.small != 0;
            grid_limit = it->second.grid;
        }
        if (h->profiling) {
            rc = launch_iteration(h, true, small, cap, grid_limit);
            if (rc) return rc;
        } else {
            hipGraphExec_t gx = nullptr;
            rc = ensure_graph(h, small, cap, grid_limit, &gx);
            if (rc) return rc;
            for (int q = 0; q < batch; q += kGraphIters) HIPCHK(h, hipGraphLaunch(gx, h->stream));
        }
        rc = read_ctrl(h, &c);
        if (rc) return rc;
        if (h->profiling) {
            for (int k = 0; k < KW_N; ++k) {
                if (c.work[k] == work_before[k]) continue;
                float ms = 0.f;
                HIPCHK(h, hipEventElapsedTime(&ms, h->ev[2 * k], h->ev[2 * k + 1]));
                h->prof_ms[k] += ms;
                h->prof_n[k] += 1;
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
        if (!h->profiling) batch = std::min(batch * 2, 64);
    }
    h->t = c.t;
    if (st) {
        st->steps = c.steps; st->rejected = c.rejected; st->rhs_evals = c.rhs_evals;
        st->cell_updates = c.cell_updates; st->active_tiles = c.n_active; st->ghost_tiles = c.n_ghost;
        st->t = c.t; st->last_dt = c.last_dt; st->min_dt = c.min_dt; st->max_dt = c.max_dt;
        st->redist_cells = c.redist_cells; st->redist_ghost = c.redist_ghost;
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
    c->sponge_strength = 0.2;
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
    if (cfg->bc < KST_BC_HALT || cfg->bc > KST_BC_SPONGE) return KST_ERR_CONFIG;
    // without periodic boundaries only interior tiles are ever active; a 1D
    // grid is one row of tiles of one row of cells
    const int min_tiles = cfg->bc == KST_BC_PERIODIC ? 1 : 3;
    const bool oned = cfg->ny_tiles * cfg->ny_per_tile == 1;
    if (cfg->nx_per_tile < 4 || cfg->nx_tiles < min_tiles) return KST_ERR_CONFIG;
    if (!oned && (cfg->ny_per_tile < 4 || cfg->ny_tiles < min_tiles)) return KST_ERR_CONFIG;
    // the 1D path runs the FP64 kernels
    if (oned && cfg->precision != KST_PREC_FP64) return KST_ERR_CONFIG;
    if (cfg->precision != KST_PREC_FP64 && cfg->precision != KST_PREC_MIXED) return KST_ERR_CONFIG;
    // mixed precision covers the hydraulic operator only
    if (cfg->precision == KST_PREC_MIXED && cfg->morphodynamics) return KST_ERR_CONFIG;
    if (cfg->n_sources > 0 && !cfg->sources) return KST_ERR_CONFIG;
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
    {
        // blocks of the 16 x 16 stage kernel the device holds at once (the
        // autotuner's resident grid)
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per, reinterpret_cast<const void*>(stage_kernel<2, false, KST_LIMITER_MINMOD2, false, B>),
                STAGE_NT, 0) != hipSuccess) {
            cus = 0; per = 0;
        }
        h->resident = 8 * ((std::max(cus * per, 8) + 7) / 8);
    }
    if (scratch_setup(h, h->device)) {
        const int e = KST_ERR_HIP;
        (void)hipStreamDestroy(h->stream);
        delete h;
        return e;
    }
    Dev& d = h->d;
    const size_t FS = (size_t)h->ntiles * h->NC;
    const size_t NV = (size_t)(h->nx + 1) * (h->ny + 1);
    d.FS = (long long)FS;
    d.morpho = cfg->morphodynamics ? 1 : 0;
    d.mixed = cfg->precision == KST_PREC_MIXED ? 1 : 0;
    rc |= dalloc(h, &d.S0, 4 * FS);
    rc |= dalloc(h, &h->vol_part, (size_t)(VOL_BLOCKS_MAX + 1) * VOL_Q * 2);
    rc |= dalloc(h, &d.UVo, 2 * FS);
    if (d.mixed) {
        // S2, S3 (4 fields) and R (5 fields) hold floats
        rc |= dalloc(h, &d.S0b, 4 * FS);
        rc |= dalloc(h, &d.S2, 2 * FS); rc |= dalloc(h, &d.S3, 2 * FS);
        rc |= dalloc(h, &d.R, (5 * FS + 1) / 2);
    } else {
        rc |= dalloc(h, &d.S2, 4 * FS); rc |= dalloc(h, &d.S3, 4 * FS);
        rc |= dalloc(h, &d.R, 5 * FS);
    }
    rc |= dalloc(h, &d.b0c, FS);
    rc |= dalloc(h, &d.xb0, (size_t)h->ntiles * h->p.NFX);
    rc |= dalloc(h, &d.yb0, (size_t)h->ntiles * h->p.NFY);
    rc |= dalloc(h, &d.vb0, (size_t)h->ntiles * NV);
    rc |= alloc_topo(h, d.TA);
    rc |= dalloc(h, &d.gsx, FS); rc |= dalloc(h, &d.gsy, FS);
    if (d.morpho) {
        rc |= alloc_topo(h, d.TB); rc |= alloc_topo(h, d.TM[0]); rc |= alloc_topo(h, d.TM[1]);
        rc |= dalloc(h, &d.SH, 4 * FS);
        for (int q = 0; q < 3; ++q) rc |= dalloc(h, &d.SM[q], 2 * FS);
        rc |= dalloc(h, &d.UV, 2 * FS);
        rc |= dalloc(h, &d.R2, 5 * FS);
        // every cell may be a candidate; the sort pads to a power of two
        size_t cap = 1;
        while (cap < FS) cap <<= 1;
        rc |= dalloc(h, &d.cand_key, cap); rc |= dalloc(h, &d.cand_ex, cap);
        rc |= dalloc(h, &d.cand_pos, FS);
        if (!rc) rc |= hipMemsetAsync(d.cand_pos, 0xFF, FS * sizeof(int), h->stream) != hipSuccess;
        d.cand_cap = (int)cap;
        d.rd_chunk = RD_THREADS;
        d.rd_lds_sort = RD_LDS;
        d.dirty_cap = 1 << 16;
        rc |= dalloc(h, &d.dirty, (size_t)d.dirty_cap);
    }
    d.act_cand = MAXCAND;
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
    // the operators' launch arguments
    d.T = d.TA;
    d.Uw = d.S0; d.Uu = d.S0 + FS; d.Uv = d.S0 + 2 * FS; d.Uh = d.S0 + 3 * FS;
    d.Sout = d.morpho ? d.SH : d.S0;
    d.second = 0;
    d.last = d.morpho ? 0 : 1;
    if (ns > 0) {
        std::vector<double> x(ns), y(ns), r(ns), tt, q, ps;
        std::vector<int> n(ns), off(ns), cells(ns);
        for (int i = 0; i < ns; ++i) {
            const kst_flux_source& s = cfg->sources[i];
            x[i] = s.x; y[i] = s.y; r[i] = s.radius; n[i] = s.n_series; cells[i] = s.num_cells;
            off[i] = (int)tt.size();
            for (int k = 0; k < s.n_series; ++k) { tt.push_back(s.time[k]); q.push_back(s.flux[k]); ps.push_back(s.psi[k]); }
            h->src_times.emplace_back(s.time, s.time + s.n_series);
        }
        (void)hipMemcpy(d.src_x, x.data(), 8 * ns, hipMemcpyHostToDevice);
        (void)hipMemcpy(d.src_y, y.data(), 8 * ns, hipMemcpyHostToDevice);
        (void)hipMemcpy(d.src_r, r.data(), 8 * ns, hipMemcpyHostToDevice);
        (void)hipMemcpy(d.src_n, n.data(), 4 * ns, hipMemcpyHostToDevice);
        (void)hipMemcpy(d.src_off, off.data(), 4 * ns, hipMemcpyHostToDevice);
        (void)hipMemcpy(d.src_cells, cells.data(), 4 * ns, hipMemcpyHostToDevice);
        (void)hipMemcpy(d.src_t, tt.data(), 8 * nk, hipMemcpyHostToDevice);
        (void)hipMemcpy(d.src_q, q.data(), 8 * nk, hipMemcpyHostToDevice);
        (void)hipMemcpy(d.src_p, ps.data(), 8 * nk, hipMemcpyHostToDevice);
    }
    h->d2 = d;
    if (d.morpho) {
        Dev& e = h->d2;
        e.T = d.TB;
        e.Uw = d.SM[2]; e.Uu = d.SH + FS; e.Uv = d.SH + 2 * FS; e.Uh = d.SM[2] + FS;
        e.R = d.R2;
        e.Sout = d.S0;
        e.second = 1;
        e.last = 1;
    }
    rc |= dalloc(h, &h->d_args, 2);
    rc |= dalloc(h, &h->d_par, 1);
    if (!rc) {
        const Dev both[2] = {h->d, h->d2};
        rc |= hipMemcpyAsync(h->d_args, both, sizeof both, hipMemcpyHostToDevice, h->stream) != hipSuccess;
        rc |= hipMemcpyAsync(h->d_par, &h->p, sizeof(Params), hipMemcpyHostToDevice, h->stream) != hipSuccess;
        rc |= hipStreamSynchronize(h->stream) != hipSuccess;
    }
    if (rc) { kst_destroy(h); return KST_ERR_HIP; }
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
    int rc = refresh_topography(h);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->have_topo = true;
    return KST_OK;
}

// NextFluxSeriesTime (src/TimeStepper.f90:281-305) on the host copy of the knots.
static double host_next_flux_time(const kst_handle* h, double t) {
    double nextT = kHuge, tdiff = kHuge;
    for (const auto& knots : h->src_times)
        for (double tt : knots) {
            const double tmp = tt - t;
            if (tmp > 0.0 && tmp < tdiff) { tdiff = tmp; nextT = tt; }
        }
    return nextT;
}

int kst_set_state(kst_handle* h, const double* w, const double* rhu, const double* rhv,
                  const double* hnpsi, const double* btv, const uint8_t* ts, double t,
                  const double* hnmax0) {
    if (!h || !w || !rhu || !rhv || !hnpsi || !ts) return KST_ERR_CONFIG;
    if (!h->have_topo) return set_err(h, KST_ERR_STATE, "kst_set_topography must come first");
    const size_t FS = (size_t)h->ntiles * h->NC;
    // tile lists: active sorted by id; ghosts = the non-active neighbours of
    // active tiles that AddGhostTiles takes (src/UpdateTiles.f90:389-481):
    // W, E, N, S and, but with periodic boundaries, the diagonal tiles (W and
    // E on a 1D grid)
    const bool periodic = h->cfg.bc == KST_BC_PERIODIC;
    auto edge = [&](int k) {
        const int gi = k % h->ntx, gj = k / h->ntx;
        return gi == 0 || gi == h->ntx - 1 || (h->nty > 1 && (gj == 0 || gj == h->nty - 1));
    };
    auto nb = [&](int k, int di, int dj) {
        int gi = k % h->ntx + di, gj = k / h->ntx + dj;
        if (periodic) {
            gi = (gi + h->ntx) % h->ntx;
            gj = (gj + h->nty) % h->nty;
        }
        return gi + gj * h->ntx;
    };
    std::vector<uint8_t> st(h->ntiles, KST_TILE_OFF);
    std::vector<int> act, gh;
    for (int k = 0; k < h->ntiles; ++k)
        if (ts[k] == KST_TILE_ACTIVE) {
            // AddTile (src/UpdateTiles.f90:56-78): an edge tile halts the run
            // with 'halt' boundaries and is left off with Dirichlet or sponge ones
            if (!periodic && edge(k)) {
                if (h->cfg.bc == KST_BC_HALT)
                    return set_err(h, KST_ERR_HALT, "Error: tried to add a tile outside the domain.");
                continue;
            }
            st[k] = KST_TILE_ACTIVE;
            act.push_back(k);
        }
    {
        const int DI[8] = {-1, 1, 0, 0, -1, 1, -1, 1}, DJ[8] = {0, 0, 1, -1, -1, -1, 1, 1};
        for (int k : act)
            for (int r = 0; r < (h->p.oned ? 2 : (periodic ? 4 : 8)); ++r) {
                const int n = nb(k, DI[r], DJ[r]);
                if (st[n] == KST_TILE_OFF) { st[n] = KST_TILE_GHOST; gh.push_back(n); }
            }
    }
    // vertex bt: each active tile takes its own copy of the global values
    // (ghost tiles keep bt = 0, as AllocateTopographicData leaves them),
    // then the topography of every set follows
    {
        const int nx = h->nx, ny = h->ny, NV1 = nx + 1;
        const size_t NV = (size_t)(nx + 1) * (ny + 1);
        std::vector<double> tv((size_t)h->ntiles * NV, 0.0);
        if (btv)
            for (int k : act) {
                const int ti = k % h->ntx, tj = k / h->ntx;
                double* dst = &tv[(size_t)k * NV];
                for (int j = 0; j <= ny; ++j)
                    for (int i = 0; i <= nx; ++i)
                        dst[j * NV1 + i] = btv[(size_t)(tj * ny + j) * (h->NX + 1) + ti * nx + i];
            }
        Topo* sets[4] = {&h->d.TA, &h->d.TB, &h->d.TM[0], &h->d.TM[1]};
        for (int q = 0; q < (h->d.morpho ? 4 : 1); ++q)
            HIPCHK(h, hipMemcpy(sets[q]->vbt, tv.data(), tv.size() * 8, hipMemcpyHostToDevice));
        int rc = refresh_topography(h);
        if (rc) return rc;
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    // fields: active tiles take the given state, ghosts the data of
    // UpdateGhostTileBCs (src/UpdateTiles.f90:571-585): on the domain edge
    // the Dirichlet boundary data, or none with periodic boundaries
    // (AllocateU's zeros), else SetDefaultTileData's dry data
    std::vector<double> b0c(FS), btc, cgam;
    HIPCHK(h, hipMemcpy(b0c.data(), h->d.b0c, FS * 8, hipMemcpyDeviceToHost));
    if (h->cfg.bc == KST_BC_DIRICHLET) {
        btc.resize(FS);
        cgam.resize(FS);
        HIPCHK(h, hipMemcpy(btc.data(), h->d.T.btc, FS * 8, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(cgam.data(), h->d.T.cgam, FS * 8, hipMemcpyDeviceToHost));
    }
    std::vector<double> S(4 * FS, 0.0);
    to_tiles(h, w, S.data());
    to_tiles(h, rhu, S.data() + FS);
    to_tiles(h, rhv, S.data() + 2 * FS);
    to_tiles(h, hnpsi, S.data() + 3 * FS);
    const kst_config& cf = h->cfg;
    const double bc_rho = cf.rhow + (cf.rhos - cf.rhow) * cf.bc_psi;
    for (int k = 0; k < h->ntiles; ++k) {
        if (st[k] == KST_TILE_ACTIVE) continue;
        const bool on_edge = st[k] == KST_TILE_GHOST && edge(k);
        for (int c = 0; c < h->NC; ++c) {
            const size_t gi = (size_t)k * h->NC + c;
            if (on_edge && cf.bc == KST_BC_PERIODIC) {
                S[gi] = 0.0;
                S[FS + gi] = 0.0; S[2 * FS + gi] = 0.0; S[3 * FS + gi] = 0.0;
            } else if (on_edge && cf.bc == KST_BC_DIRICHLET) {
                const double hn = cf.bc_hn;
                double wv = btc[gi] + hn / cgam[gi];
                wv = wv + b0c[gi];
                S[gi] = wv;
                S[FS + gi] = bc_rho * hn * cf.bc_u;
                S[2 * FS + gi] = bc_rho * hn * cf.bc_v;
                S[3 * FS + gi] = hn * cf.bc_psi;
            } else {
                S[gi] = b0c[gi];
                S[FS + gi] = 0.0; S[2 * FS + gi] = 0.0; S[3 * FS + gi] = 0.0;
            }
        }
    }
    HIPCHK(h, hipMemcpy(h->d.S0, S.data(), S.size() * 8, hipMemcpyHostToDevice));
    if (h->d.mixed) {
        HIPCHK(h, hipMemcpy(h->d.S0b, S.data(), S.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemset(h->d.S2, 0, 4 * FS * 4));
        HIPCHK(h, hipMemset(h->d.S3, 0, 4 * FS * 4));
        HIPCHK(h, hipMemset(h->d.R, 0, 5 * FS * 4));
    } else {
        HIPCHK(h, hipMemcpy(h->d.S2, S.data(), S.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->d.S3, S.data(), S.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemset(h->d.R, 0, 5 * FS * 8));
    }
    if (h->d.morpho) {
        HIPCHK(h, hipMemcpy(h->d.SH, S.data(), S.size() * 8, hipMemcpyHostToDevice));
        for (int q = 0; q < 3; ++q) {
            HIPCHK(h, hipMemcpy(h->d.SM[q], S.data(), FS * 8, hipMemcpyHostToDevice));
            HIPCHK(h, hipMemcpy(h->d.SM[q] + FS, S.data() + 3 * FS, FS * 8, hipMemcpyHostToDevice));
        }
        HIPCHK(h, hipMemset(h->d.UV, 0, 2 * FS * 8));
        HIPCHK(h, hipMemset(h->d.R2, 0, 5 * FS * 8));
    }
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
        (void)hipMemcpy(&sx, h->d.src_x + q, 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(&sy, h->d.src_y + q, 8, hipMemcpyDeviceToHost);
        (void)hipMemcpy(&sr, h->d.src_r + q, 8, hipMemcpyDeviceToHost);
        for (int J = 0; J < h->NY; ++J)
            for (int I = 0; I < h->NX; ++I) {
                const int ti = I / h->nx, tj = J / h->ny, i = I % h->nx, j = J % h->ny;
                const double x = -0.5 * h->cfg.x_size + h->cfg.dx * ((ti + 1 - 1.0) * h->nx + (i + 1 - 0.5));
                const double y = -0.5 * h->cfg.y_size + h->cfg.dy * ((tj + 1 - 1.0) * h->ny + (j + 1 - 0.5));
                // along x only on a 1D grid (src/SetSources.f90:219-231)
                const double r2 = h->p.oned ? (x - sx) * (x - sx) : (x - sx) * (x - sx) + (y - sy) * (y - sy);
                if (r2 <= sr * sr) hs[ti + tj * h->ntx] = 1;
            }
    }
    HIPCHK(h, hipMemcpy(h->d.has_src, hs.data(), hs.size(), hipMemcpyHostToDevice));
    Ctrl c;
    std::memset(&c, 0, sizeof c);
    c.t = t;
    c.tend = t;
    c.next_t = host_next_flux_time(h, t);
    const unsigned long long huge = 0x7FEFFFFFFFFFFFFFull;
    for (int k = 0; k < 6; ++k) c.slot[k].bits = huge;
    c.status = ST_DONE;
    c.n_active = (int)act.size();
    c.n_ghost = (int)gh.size();
    int rc = write_ctrl(h, &c);
    if (rc) return rc;
    if (!gh.empty())
        hipLaunchKernelGGL(init_ghosts_kernel, dim3(gh.size()), dim3(256), 0, h->stream, h->d, h->p);
    if (!act.empty())
        hipLaunchKernelGGL(init_flags_kernel, dim3(act.size()), dim3(256), 0, h->stream, h->d, h->p);
    // tileContainer's velocities after the state is set: desingularised from it
    HIPCHK(h, hipMemsetAsync(h->d.UVo, 0, 2 * (size_t)h->ntiles * h->NC * sizeof(double), h->stream));
    if (!act.empty())
        hipLaunchKernelGGL(uv_init_kernel, dim3(std::min((int)act.size(), VOL_BLOCKS_MAX)), dim3(VOL_NT), 0,
                           h->stream, h->d_args, h->d_par);
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
    c.redist_cells = c.redist_ghost = 0;
    c.mref = 0; c.ncand = 0; c.commit = 0;
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
    const double* Sn = h->d.S0;
    if (h->d.mixed) {
        Ctrl c;
        int rc = read_ctrl(h, &c);
        if (rc) return rc;
        if (c.cur) Sn = h->d.S0b;
    }
    HIPCHK(h, hipMemcpy(S.data(), Sn, S.size() * 8, hipMemcpyDeviceToHost));
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
        std::vector<double> b0c(FS), bt(FS), g(FS), H(FS, 0.0);
        HIPCHK(h, hipMemcpy(b0c.data(), h->d.b0c, FS * 8, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(bt.data(), h->d.TA.btc, FS * 8, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(g.data(), h->d.TA.cgam, FS * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < FS; ++i) {
            const int k = (int)(i / h->NC);
            if (st[k] == KST_TILE_OFF) continue;
            if (st[k] == KST_TILE_GHOST) {
                // a ghost's Hn is the one it was given: zero, or the
                // Dirichlet boundary depth on the domain edge
                const int gi = k % h->ntx, gj = k / h->ntx;
                const bool edge = gi == 0 || gi == h->ntx - 1 || (h->nty > 1 && (gj == 0 || gj == h->nty - 1));
                H[i] = (h->cfg.bc == KST_BC_DIRICHLET && edge) ? h->cfg.bc_hn : 0.0;
                continue;
            }
            double x = -bt[i];
            x = x + (S[i] - b0c[i]);
            x = x * g[i];
            H[i] = x < 0.0 ? 0.0 : x;
        }
        from_tiles(h, H.data(), hn);
    }
    if (btv) {
        // vertex (I, J) from the copy of the tile holding cell (min(I, NX-1), min(J, NY-1))
        const int nx = h->nx, ny = h->ny, NV1 = nx + 1;
        const size_t NV = (size_t)(nx + 1) * (ny + 1);
        std::vector<double> tv((size_t)h->ntiles * NV);
        HIPCHK(h, hipMemcpy(tv.data(), h->d.TA.vbt, tv.size() * 8, hipMemcpyDeviceToHost));
        for (int J = 0; J <= h->NY; ++J)
            for (int I = 0; I <= h->NX; ++I) {
                const int ci = std::min(I, h->NX - 1), cj = std::min(J, h->NY - 1);
                const int k = (cj / ny) * h->ntx + ci / nx;
                double v = 0.0;
                if (st[k] != KST_TILE_OFF) v = tv[(size_t)k * NV + (J - (cj / ny) * ny) * NV1 + (I - (ci / nx) * nx)];
                btv[(size_t)J * (h->NX + 1) + I] = v;
            }
    }
    if (ts) std::memcpy(ts, st.data(), st.size());
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    if (t) *t = c.t;
    return KST_OK;
}

namespace {

// CalculateVolume with the partial sums in `part`, which holds `cap` doubles
// (VOL_BLOCKS_MAX + 1 groups of VOL_Q pairs suffice for any active set).
int volume_into(kst_handle* h, double* part, size_t cap, double out[6]) {
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    const int nb = std::max(1, std::min(c.n_active, VOL_BLOCKS_MAX));
    if ((size_t)(VOL_BLOCKS_MAX + 1) * VOL_Q * 2 > cap)
        return set_err(h, KST_ERR_CONFIG, "volume: the partial-sum buffer holds " + std::to_string(cap) +
                                              " doubles; " + std::to_string((VOL_BLOCKS_MAX + 1) * VOL_Q * 2) +
                                              " are needed");
    double* fin = part + (size_t)VOL_BLOCKS_MAX * VOL_Q * 2;
    hipLaunchKernelGGL(volume_kernel, dim3(nb), dim3(VOL_NT), 0, h->stream, h->d_args, h->d_par, part);
    hipLaunchKernelGGL(volume_final_kernel, dim3(1), dim3(VOL_NT), 0, h->stream, part, nb, fin);
    HIPCHK(h, hipGetLastError());
    double r[VOL_Q * 2];
    HIPCHK(h, hipMemcpyAsync(r, fin, sizeof r, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const kst_config& g = h->cfg;
    // each sum times dx dy, then the columns as CalculateVolume writes them
    // (src/Output.f90:709-730)
    double v[VOL_Q];
    for (int q = 0; q < VOL_Q; ++q) v[q] = (r[2 * q] + r[2 * q + 1]) * g.dx * g.dy;
    const double vol = v[0], mass = v[1], bed = v[2], solids = v[3];
    const double rhob = g.rhow * g.bed_porosity + g.rhos * (1.0 - g.bed_porosity);
    out[0] = vol;
    out[1] = bed;
    out[2] = mass;
    out[3] = bed * rhob;
    out[4] = solids * g.rhos;
    out[5] = bed * g.rhos * (1.0 - g.bed_porosity);
    return KST_OK;
}

}  // namespace

int kst_volume(kst_handle* h, double out[6]) {
    if (!h || !out) return KST_ERR_CONFIG;
    if (!h->have_state) return set_err(h, KST_ERR_STATE, "kst_set_state must come first");
    return volume_into(h, h->vol_part, (size_t)(VOL_BLOCKS_MAX + 1) * VOL_Q * 2, out);
}

int kst_volume_noalloc(kst_handle* h, double out[6]) {
    if (!h || !out) return KST_ERR_CONFIG;
    if (!h->have_state) return set_err(h, KST_ERR_STATE, "kst_set_state must come first");
    int cur = -1;
    HIPCHK(h, hipGetDevice(&cur));
    if (cur != h->device)
        return set_err(h, KST_ERR_CONFIG, "kst_volume_noalloc: the handle's device must be current");
    double* part = scratch_ptr(h->device);
    if (!part) return set_err(h, KST_ERR_STATE, "kst_volume_noalloc: no scratch address for this device");
    // the volume kernels write at most VOL_BLOCKS_MAX partial groups and one
    // merged group, which is what scratch holds; volume_into refuses a
    // buffer smaller than that
    return volume_into(h, part, kScratchDoubles, out);
}

int kst_get_output(kst_handle* h, int32_t max_tiles, double* fields, int32_t* tiles, int32_t* n_tiles) {
    if (!h || !fields || !tiles || !n_tiles) return KST_ERR_CONFIG;
    if (!h->have_state) return set_err(h, KST_ERR_STATE, "kst_set_state must come first");
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    const int na = c.n_active;
    if (na > max_tiles)
        return set_err(h, KST_ERR_CONFIG, "kst_get_output: max_tiles is smaller than the active set (" +
                                              std::to_string(na) + " tiles)");
    *n_tiles = na;
    if (na == 0) return KST_OK;
    const size_t need = (size_t)na * OUT_COLS * h->NC;
    if (need > h->out_cap) {
        if (h->out_buf) {
            HIPCHK(h, hipStreamSynchronize(h->stream));
            HIPCHK(h, hipFree(h->out_buf));
            h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), (void*)h->out_buf), h->allocs.end());
            h->out_buf = nullptr;
            h->out_cap = 0;
        }
        rc = dalloc(h, &h->out_buf, need);
        if (rc) return rc;
        h->out_cap = need;
    }
    const int nb = std::min(na, VOL_BLOCKS_MAX);
    hipLaunchKernelGGL(output_kernel, dim3(nb), dim3(VOL_NT), 0, h->stream, h->d_args, h->d_par, h->out_buf);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(fields, h->out_buf, need * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(tiles, h->d.active, (size_t)na * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
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

int kst_set_maxima(kst_handle* h, const double* const mx[11]) {
    if (!h || !mx) return KST_ERR_CONFIG;
    const size_t FS = (size_t)h->ntiles * h->NC;
    std::vector<double> M(11 * FS);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(M.data(), h->d.maxima, M.size() * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < 11; ++q)
        if (mx[q]) to_tiles(h, mx[q], M.data() + q * FS);
    HIPCHK(h, hipMemcpy(h->d.maxima, M.data(), M.size() * 8, hipMemcpyHostToDevice));
    return KST_OK;
}

int kst_get_tile_bt(kst_handle* h, double* bt_tiles) {
    if (!h || !bt_tiles) return KST_ERR_CONFIG;
    const size_t n = (size_t)(h->nx + 1) * (h->ny + 1) * h->ntiles;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(bt_tiles, h->d.TA.vbt, n * 8, hipMemcpyDeviceToHost));
    return KST_OK;
}

int kst_set_cell_topography(kst_handle* h, const double* btc, const double* dbdx, const double* dbdy,
                            const uint8_t* cells) {
    if (!h || !btc || !dbdx || !dbdy || !cells) return KST_ERR_CONFIG;
    if (!h->have_state) return set_err(h, KST_ERR_STATE, "kst_set_state must come first");
    const size_t FS = (size_t)h->ntiles * h->NC, NC = (size_t)h->NX * h->NY;
    std::vector<double> g(NC), in(FS), cur(FS);
    std::vector<double> mk(NC), mt(FS);
    for (size_t q = 0; q < NC; ++q) mk[q] = cells[q] ? 1.0 : 0.0;
    to_tiles(h, mk.data(), mt.data());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    Topo* sets[4] = {&h->d.TA, &h->d.TB, &h->d.TM[0], &h->d.TM[1]};
    const int nsets = h->d.morpho ? 4 : 1;
    const double* src[3] = {btc, dbdx, dbdy};
    for (int f = 0; f < 3; ++f) {
        double* dst = f == 0 ? h->d.TA.btc : (f == 1 ? h->d.TA.cbx : h->d.TA.cby);
        HIPCHK(h, hipMemcpy(cur.data(), dst, FS * 8, hipMemcpyDeviceToHost));
        to_tiles(h, src[f], in.data());
        for (size_t q = 0; q < FS; ++q)
            if (mt[q] != 0.0) cur[q] = in[q];
        for (int k = 0; k < nsets; ++k) {
            double* to = f == 0 ? sets[k]->btc : (f == 1 ? sets[k]->cbx : sets[k]->cby);
            HIPCHK(h, hipMemcpy(to, cur.data(), FS * 8, hipMemcpyHostToDevice));
        }
    }
    for (int k = 0; k < nsets; ++k)
        hipLaunchKernelGGL(cell_gamma_kernel, dim3(h->ntiles), dim3(256), 0, h->stream, h->d, h->p, *sets[k]);
    for (int k = 0; k < nsets; ++k)
        hipLaunchKernelGGL(topo2_kernel, dim3(h->ntiles), dim3(256), 0, h->stream, h->d, h->p, *sets[k]);
    // what kst_set_state derived from the cell topography: the activation
    // flags of the active tiles and tileContainer's velocities
    Ctrl c;
    int rc = read_ctrl(h, &c);
    if (rc) return rc;
    HIPCHK(h, hipMemset(h->d.eflags, 0, (size_t)h->ntiles * 4));
    if (c.n_active > 0)
        hipLaunchKernelGGL(init_flags_kernel, dim3(c.n_active), dim3(256), 0, h->stream, h->d, h->p);
    HIPCHK(h, hipMemsetAsync(h->d.UVo, 0, 2 * FS * sizeof(double), h->stream));
    if (c.n_active > 0)
        hipLaunchKernelGGL(uv_init_kernel, dim3(std::min(c.n_active, VOL_BLOCKS_MAX)), dim3(VOL_NT), 0,
                           h->stream, h->d_args, h->d_par);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return KST_OK;
}

int kst_get_cell_topography(kst_handle* h, double* b0c, double* btc, double* dbdx, double* dbdy) {
    if (!h) return KST_ERR_CONFIG;
    const size_t FS = (size_t)h->ntiles * h->NC;
    std::vector<double> v(FS);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double* outs[4] = {b0c, btc, dbdx, dbdy};
    const double* srcs[4] = {h->d.b0c, h->d.TA.btc, h->d.TA.cbx, h->d.TA.cby};
    for (int q = 0; q < 4; ++q) {
        if (!outs[q]) continue;
        HIPCHK(h, hipMemcpy(v.data(), srcs[q], FS * 8, hipMemcpyDeviceToHost));
        from_tiles(h, v.data(), outs[q]);
    }
    return KST_OK;
}

int kst_last_error(const kst_handle* h, char* buf, int buflen) {
    if (!h || !buf || buflen <= 0) return KST_ERR_CONFIG;
    std::snprintf(buf, buflen, "%s", h->err.c_str());
    return KST_OK;
}

int kst_set_grid_cap(kst_handle* h, int tiles) {
    if (!h || tiles < 0) return KST_ERR_CONFIG;
    h->fixed_cap = tiles;
    return KST_OK;
}

int kst_autotune(kst_handle* h, int enable) {
    if (!h) return KST_ERR_CONFIG;
    h->autotune = enable != 0;
    h->tuned.clear();
    return KST_OK;
}

int kst_autotune_choices(const kst_handle* h, int max, int32_t* bucket, int32_t* small,
                         int32_t* grid, double* us) {
    if (!h || max < 0) return KST_ERR_CONFIG;
    int n = 0;
    for (const auto& kv : h->tuned) {
        if (n >= max) break;
        if (bucket) bucket[n] = kv.first;
        if (small) small[n] = kv.second.small;
        if (grid) grid[n] = kv.second.grid;
        if (us)
            for (int q = 0; q < 4; ++q) us[4 * n + q] = kv.second.us[q];
        ++n;
    }
    return n;
}

int kst_set_redistribution_chunk(kst_handle* h, int chunk, int lds_sort) {
    if (!h) return KST_ERR_CONFIG;
    if (chunk < 0 || chunk > RD_THREADS || lds_sort < 0 || lds_sort > RD_LDS)
        return set_err(h, KST_ERR_CONFIG, "kst_set_redistribution_chunk: chunk must lie in 1.." +
                                              std::to_string(RD_THREADS) + " and lds_sort in 1.." +
                                              std::to_string(RD_LDS) + " (0: default)");
    if (!h->d.morpho) return KST_OK;
    const int c = chunk ? chunk : RD_THREADS, l = lds_sort ? lds_sort : RD_LDS;
    h->d.rd_chunk = h->d2.rd_chunk = c;
    h->d.rd_lds_sort = h->d2.rd_lds_sort = l;
    const Dev both[2] = {h->d, h->d2};
    HIPCHK(h, hipMemcpyAsync(h->d_args, both, sizeof both, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return KST_OK;
}

int kst_set_activation_candidates(kst_handle* h, int cap) {
    if (!h) return KST_ERR_CONFIG;
    if (cap < 0 || cap > MAXCAND)
        return set_err(h, KST_ERR_CONFIG, "kst_set_activation_candidates: cap must lie in 1.." +
                                              std::to_string(MAXCAND) + " (0: default)");
    h->d.act_cand = h->d2.act_cand = cap ? cap : MAXCAND;
    const Dev both[2] = {h->d, h->d2};
    HIPCHK(h, hipMemcpyAsync(h->d_args, both, sizeof both, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return KST_OK;
}

int kst_profile(kst_handle* h, int enable) {
    if (!h) return KST_ERR_CONFIG;
    h->profiling = enable != 0;
    for (int k = 0; k < KW_N; ++k) { h->prof_ms[k] = 0.0; h->prof_n[k] = 0; }
    return KST_OK;
}

int kst_profile_read(kst_handle* h, double ms[KST_PROFILE_KERNELS], int64_t launches[KST_PROFILE_KERNELS]) {
    if (!h) return KST_ERR_CONFIG;
    for (int k = 0; k < KW_N; ++k) {
        if (ms) ms[k] = h->prof_n[k] ? h->prof_ms[k] / h->prof_n[k] : 0.0;
        if (launches) launches[k] = h->prof_n[k];
        h->prof_ms[k] = 0.0;
        h->prof_n[k] = 0;
    }
    return KST_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------

static int batch_err(kst_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

#define BCHK(b, call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return batch_err(b, KST_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBatchIters = 8;

typedef void (*BatchStageFn)(Batch, int);
struct BatchSet { BatchStageFn fn[4]; };

// MinMod2 (the reference's default) at compile time, or any limiter per
// scenario (LIM_RUNTIME), once a scenario with another one joins
template <bool VISC, bool MIX, int L>
BatchSet make_batch_set() {
    return BatchSet{{batch_stage_kernel<1, VISC, L, MIX>, batch_stage_kernel<2, VISC, L, MIX>,
                     batch_stage_kernel<3, VISC, L, MIX>, batch_stage_kernel<4, VISC, L, MIX>}};
}

static const BatchSet& batch_set(bool visc, bool mixed, bool anylim) {
    constexpr int M2 = KST_LIMITER_MINMOD2, RT = LIM_RUNTIME;
    static const BatchSet sets[2][2][2] = {
        {{make_batch_set<false, false, M2>(), make_batch_set<false, true, M2>()},
         {make_batch_set<true, false, M2>(), make_batch_set<true, true, M2>()}},
        {{make_batch_set<false, false, RT>(), make_batch_set<false, true, RT>()},
         {make_batch_set<true, false, RT>(), make_batch_set<true, true, RT>()}}};
    return sets[anylim ? 1 : 0][visc ? 1 : 0][mixed ? 1 : 0];
}

// One iteration: activation, the work lists, the first hydraulic operator;
// with morphodynamic slots their M(2dt) (three stages, redistribution, face
// topography of the new bed) and second hydraulic operator; step control;
// the accepted beds' commit (the launch sequence of one scenario).
static int batch_launch_iteration(kst_batch* b) {
    const Batch& bt = b->bt;
    const BatchSet& k = batch_set(b->visc == 1, b->mixed == 1, b->anylim == 1);
    const bool mo = b->morpho != 0;
    hipLaunchKernelGGL(batch_activate_kernel, dim3(b->nslot), dim3(ACT_NT), 0, b->stream, bt);
    hipLaunchKernelGGL(batch_plan_kernel, dim3(1), dim3(BATCH_PLAN_THREADS), 0, b->stream, bt, mo ? 3 : 1);
    for (int q = 0; q < 4; ++q)
        hipLaunchKernelGGL(k.fn[q], dim3(b->grid), dim3(STAGE_NT), 0, b->stream, bt, 0);
    if (mo) {
        constexpr int NT = morpho_nt<B>();
        hipLaunchKernelGGL((batch_morpho_kernel<1>), dim3(b->mgrid), dim3(NT), 0, b->stream, bt);
        hipLaunchKernelGGL((batch_morpho_kernel<2>), dim3(b->mgrid), dim3(NT), 0, b->stream, bt);
        hipLaunchKernelGGL((batch_morpho_kernel<3>), dim3(b->mgrid), dim3(NT), 0, b->stream, bt);
        hipLaunchKernelGGL(batch_redistribute_kernel, dim3(b->nslot), dim3(RD_THREADS), 0, b->stream, bt);
        hipLaunchKernelGGL((batch_tile_part_kernel<0>), dim3(b->tgrid), dim3(256), 0, b->stream, bt);
        for (int q = 0; q < 4; ++q)
            hipLaunchKernelGGL(k.fn[q], dim3(b->grid), dim3(STAGE_NT), 0, b->stream, bt, 1);
    }
    hipLaunchKernelGGL(batch_finalize_kernel, dim3((b->nslot + 63) / 64), dim3(64), 0, b->stream, bt);
    if (mo) hipLaunchKernelGGL((batch_tile_part_kernel<1>), dim3(b->tgrid), dim3(256), 0, b->stream, bt);
    BCHK(b, hipGetLastError());
    return 0;
}

static int batch_ensure_graph(kst_batch* b) {
    if (b->gexec) return 0;
    hipGraph_t g = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int q = 0; q < kBatchIters; ++q) {
        int rc = batch_launch_iteration(b);
        if (rc) {
            hipGraph_t dummy = nullptr;
            (void)hipStreamEndCapture(b->stream, &dummy);
            if (dummy) (void)hipGraphDestroy(dummy);
            return rc;
        }
    }
    BCHK(b, hipStreamEndCapture(b->stream, &g));
    BCHK(b, hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    BCHK(b, hipGraphDestroy(g));
    return 0;
}

static int slot_state(int st) {
    if (st == ST_NEW || st == ST_RETRY) return KST_SLOT_RUNNING;
    if (st == ST_DONE) return KST_SLOT_DONE;
    return KST_SLOT_ERROR;
}

extern "C" {

int kst_batch_create(int max_slots, kst_batch** out) {
    if (!out || max_slots <= 0 || max_slots > 65536) return KST_ERR_CONFIG;
    *out = nullptr;
    kst_batch* b = new kst_batch();
    b->nslot = max_slots;
    b->slots.assign(max_slots, nullptr);
    b->on.assign(max_slots, 0);
    auto fail = [&](int code) { kst_batch_destroy(b); return code; };
    if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess)
        return fail(KST_ERR_HIP);
    if (hipMalloc(&b->d_dev, sizeof(Dev) * max_slots) != hipSuccess ||
        hipMalloc(&b->d_dev2, sizeof(Dev) * max_slots) != hipSuccess ||
        hipMalloc(&b->d_par, sizeof(Params) * max_slots) != hipSuccess ||
        hipMalloc(&b->d_on, sizeof(int) * max_slots) != hipSuccess ||
        hipMalloc(&b->d_off, sizeof(int) * (max_slots + 1)) != hipSuccess ||
        hipMalloc(&b->d_off2, sizeof(int) * (max_slots + 1)) != hipSuccess ||
        hipMalloc(&b->d_off3, sizeof(int) * (max_slots + 1)) != hipSuccess ||
        hipMalloc(&b->d_status, sizeof(int) * max_slots) != hipSuccess)
        return fail(KST_ERR_HIP);
    if (hipMemset(b->d_on, 0, sizeof(int) * max_slots) != hipSuccess ||
        hipMemset(b->d_off, 0, sizeof(int) * (max_slots + 1)) != hipSuccess ||
        hipMemset(b->d_off2, 0, sizeof(int) * (max_slots + 1)) != hipSuccess ||
        hipMemset(b->d_off3, 0, sizeof(int) * (max_slots + 1)) != hipSuccess ||
        hipMemset(b->d_status, 0, sizeof(int) * max_slots) != hipSuccess)
        return fail(KST_ERR_HIP);
    b->bt = Batch{b->d_dev, b->d_par, b->d_on, b->d_off, b->d_status, max_slots,
                  b->d_dev2, b->d_off2, b->d_off3};
    // resident grids: each kernel's occupancy on every CU, a multiple of 8
    int cus = 0, per_cu = 0, per_cu_m = 0, per_cu_t = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, batch_stage_kernel<2, true, KST_LIMITER_MINMOD2, false>, STAGE_NT, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_m, batch_morpho_kernel<2>, morpho_nt<B>(), 0) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_t, batch_tile_part_kernel<0>, 256, 0) != hipSuccess)
        return fail(KST_ERR_HIP);
    b->grid = std::max(8, (cus * std::max(per_cu, 1)) / 8 * 8);
    b->mgrid = std::max(8, (cus * std::max(per_cu_m, 1)) / 8 * 8);
    b->tgrid = std::max(8, (cus * std::max(per_cu_t, 1)) / 8 * 8);
    *out = b;
    return KST_OK;
}

int kst_batch_attach(kst_batch* b, kst_handle* h, int* slot) {
    if (!b || !h || !slot) return KST_ERR_CONFIG;
    if (!h->have_state) return batch_err(b, KST_ERR_STATE, "kst_set_state must come before kst_batch_attach");
    if (h->p.oned) return batch_err(b, KST_ERR_CONFIG, "batches run 2D grids only");
    const int visc = h->p.nu != 0.0 ? 1 : 0, mixed = h->d.mixed ? 1 : 0;
    bool any = false;
    for (int q = 0; q < b->nslot; ++q) any = any || b->on[q];
    if (any && (visc != b->visc || mixed != b->mixed))
        return batch_err(b, KST_ERR_CONFIG, "a batch holds scenarios of one precision and one eddy-viscosity setting");
    int s = -1;
    for (int q = 0; q < b->nslot && s < 0; ++q)
        if (!b->slots[q]) s = q;
    if (s < 0) return batch_err(b, KST_ERR_CONFIG, "no free slot");
    const int anylim = ((any && b->anylim) || h->p.limiter != KST_LIMITER_MINMOD2) ? 1 : 0;
    const int morpho = ((any && b->morpho) || h->d.morpho) ? 1 : 0;
    if (b->visc != visc || b->mixed != mixed || b->anylim != anylim || b->morpho != morpho) {
        // the graph's kernels depend on the setting
        if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
        b->visc = visc;
        b->mixed = mixed;
        b->anylim = anylim;
        b->morpho = morpho;
    }
    BCHK(b, hipStreamSynchronize(h->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    const int one = 1, st = ST_DONE;
    BCHK(b, hipMemcpy(b->d_dev + s, &h->d, sizeof(Dev), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->d_dev2 + s, &h->d2, sizeof(Dev), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->d_par + s, &h->p, sizeof(Params), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->d_status + s, &st, sizeof(int), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->d_on + s, &one, sizeof(int), hipMemcpyHostToDevice));
    b->slots[s] = h;
    b->on[s] = 1;
    *slot = s;
    return KST_OK;
}

int kst_batch_detach(kst_batch* b, int slot) {
    if (!b || slot < 0 || slot >= b->nslot || !b->slots[slot]) return KST_ERR_CONFIG;
    BCHK(b, hipStreamSynchronize(b->stream));
    const int zero = 0;
    BCHK(b, hipMemcpy(b->d_on + slot, &zero, sizeof(int), hipMemcpyHostToDevice));
    b->slots[slot] = nullptr;
    b->on[slot] = 0;
    return KST_OK;
}

int kst_batch_start(kst_batch* b, int slot, double tend, int64_t nsteps) {
    if (!b || slot < 0 || slot >= b->nslot || !b->slots[slot]) return KST_ERR_CONFIG;
    kst_handle* h = b->slots[slot];
    BCHK(b, hipStreamSynchronize(b->stream));
    Ctrl c;
    if (read_ctrl(h, &c)) return batch_err(b, KST_ERR_HIP, h->err);
    if (c.status == ST_ERROR) return batch_err(b, c.err, "the scenario is in an error state");
    const bool by_steps = nsteps > 0;
    if (!by_steps && !(tend > c.t)) return batch_err(b, KST_ERR_CONFIG, "tend must lie after the scenario's time");
    c.status = ST_NEW;
    c.mode = by_steps ? 1 : 0;
    c.tend = by_steps ? DBL_MAX : tend;
    c.steps_left = by_steps ? nsteps : 0;
    c.steps = c.rejected = c.rhs_evals = c.cell_updates = 0;
    c.redist_cells = c.redist_ghost = 0;
    c.mref = 0; c.ncand = 0; c.commit = 0;
    c.min_dt = 0.0; c.max_dt = 0.0;
    if (write_ctrl(h, &c)) return batch_err(b, KST_ERR_HIP, h->err);
    const int st = ST_NEW;
    BCHK(b, hipMemcpy(b->d_status + slot, &st, sizeof(int), hipMemcpyHostToDevice));
    return KST_OK;
}

int kst_batch_run(kst_batch* b, int64_t max_iters, int* status) {
    if (!b) return KST_ERR_CONFIG;
    std::vector<int> st(b->nslot, ST_DONE);
    int64_t iters = 0;
    bool running = false;
    for (int q = 0; q < b->nslot; ++q) running = running || b->on[q];
    if (running) {
        int rc = batch_ensure_graph(b);
        if (rc) return rc;
    }
    while (running && iters < max_iters) {
        BCHK(b, hipGraphLaunch(b->gexec, b->stream));
        iters += kBatchIters;
        BCHK(b, hipMemcpyAsync(st.data(), b->d_status, sizeof(int) * b->nslot, hipMemcpyDeviceToHost, b->stream));
        BCHK(b, hipStreamSynchronize(b->stream));
        running = false;
        for (int q = 0; q < b->nslot; ++q)
            if (b->on[q] && (st[q] == ST_NEW || st[q] == ST_RETRY)) running = true;
    }
    if (!running) {
        BCHK(b, hipMemcpyAsync(st.data(), b->d_status, sizeof(int) * b->nslot, hipMemcpyDeviceToHost, b->stream));
        BCHK(b, hipStreamSynchronize(b->stream));
    }
    if (status)
        for (int q = 0; q < b->nslot; ++q) status[q] = b->on[q] ? slot_state(st[q]) : KST_SLOT_FREE;
    return KST_OK;
}

int kst_batch_stats(kst_batch* b, int slot, kst_stats* out) {
    if (!b || !out || slot < 0 || slot >= b->nslot || !b->slots[slot]) return KST_ERR_CONFIG;
    kst_handle* h = b->slots[slot];
    BCHK(b, hipStreamSynchronize(b->stream));
    Ctrl c;
    if (read_ctrl(h, &c)) return batch_err(b, KST_ERR_HIP, h->err);
    std::memset(out, 0, sizeof *out);
    out->steps = c.steps; out->rejected = c.rejected; out->rhs_evals = c.rhs_evals;
    out->cell_updates = c.cell_updates; out->active_tiles = c.n_active; out->ghost_tiles = c.n_ghost;
    out->t = c.t; out->last_dt = c.last_dt; out->min_dt = c.min_dt; out->max_dt = c.max_dt;
    h->t = c.t;
    if (c.status == ST_ERROR) {
        if (c.err == KST_ERR_HALT)
            return set_err(h, KST_ERR_HALT, "Error: tried to add a tile outside the domain. Try increasing the "
                                            "domain size or repositioning its location.");
        return set_err(h, c.err ? c.err : KST_ERR_STATE, "device error");
    }
    return KST_OK;
}

int kst_batch_last_error(const kst_batch* b, char* buf, int buflen) {
    if (!b || !buf || buflen <= 0) return KST_ERR_CONFIG;
    std::snprintf(buf, buflen, "%s", b->err.c_str());
    return KST_OK;
}

void kst_batch_destroy(kst_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    (void)hipFree(b->d_dev); (void)hipFree(b->d_dev2); (void)hipFree(b->d_par); (void)hipFree(b->d_on);
    (void)hipFree(b->d_off); (void)hipFree(b->d_off2); (void)hipFree(b->d_off3); (void)hipFree(b->d_status);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

void kst_destroy(kst_handle* h) {
    if (!h) return;
    for (const auto& kv : h->graphs)
        if (kv.second) (void)hipGraphExecDestroy(kv.second);
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void* p : h->allocs) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
