// kst_morpho.hpp -- gfx950 kernels of the morphodynamic operator M(2dt) of
// the Strang step H(dt) M(2dt) H(dt) (src/TimeStepper.f90:532-781) and of
// the bed bookkeeping around it.  Included by kst_step.hip after the stage
// kernels (it uses Dev, Ctrl, Topo and the work mapping defined there).
//
//   morpho_kernel<1..3>   one SSP stage on the vertex bt of every active tile:
//                         E - D with the dryness check, vertex rates, the
//                         clamped stage update, cell topography and the
//                         conservative w / Hnpsi update; stage 3 also runs
//                         the excess-deposition and depth-change checks.
//   redistribute_kernel   the sorted redistribution list (one workgroup).
//   face_topo_kernel      face topography and CFL numerators of the new bed.
//   commit_kernel         the new bed becomes current once the step is accepted.
#pragma once

namespace kst {

constexpr int MR1 = B + 2;   // sub-block plus a 1-cell halo
constexpr int MV = B + 1;    // vertices along a sub-block edge
// Threads of a morphodynamic block of 16 x BY cells: enough for every phase
// (Hn on the 20 x (BY+4) region, E - D on 18 x (BY+2), 17 x (BY+1) vertices,
// 16 x BY cells) to take one round, so the transcendental-heavy E - D is
// evaluated once on the critical path.  Small active sets use BY = 4: the
// E - D of a 16 x 16 block is ~6 waves of FP64 work on one CU, of a 16 x 4
// block 2 waves on two SIMDs, and the chip has idle CUs for the extra blocks.
template <int BY>
constexpr int morpho_nt() { return BY == B ? 512 : 256; }
static_assert(morpho_nt<B>() >= NREG && morpho_nt<B>() >= MR1 * MR1, "one round per phase");
static_assert(morpho_nt<B / 4>() >= RW * (B / 4 + 4), "one round per phase");

// Thread 0: whether the morphodynamic operator runs in this attempt (the
// first hydraulic operator passed its re-checks), its dt, and the block's
// work item (sub-block of an active tile, XCD-chunked as the stage kernel).
struct MWork {
    int tile, q, bx0, by0, bxa, bya, skip, more;
    double dt;
};

// witem >= 0: that item of the scenario's list (a batch's work list), else
// the block's pass-th item of a one-scenario launch.
template <int BY, class DevT>
__device__ void morpho_work(const DevT& d, const Params& p, MWork& m, bool need_no_refine, int kw,
                            int pass, int witem = -1) {
    const int sbx = (p.nx + B - 1) / B, sby = (p.ny + BY - 1) / BY, spt = sbx * sby;
    // the small shape's item and tile ahead of the control block (cap_item)
    constexpr bool SPEC = BY < B;
    int wspec = 0, tspec = 0;
    if (SPEC && witem < 0) {
        wspec = cap_item();
        tspec = d.active[min(wspec / spt, p.ntiles)];   // the list holds ntiles + 1 entries
    }
    const CtrlSnap cs = snap_ctrl(ctrl_view(d));
    const CtrlSnap* c = &cs;
    int skip = (c->status != ST_NEW && c->status != ST_RETRY);
    double dt = 0.0;
    if (!skip) {
        dt = step_dt(d, p, c);
        if (!first_operator_ok(c, p, dt)) skip = 1;
        if (need_no_refine && c->mref) skip = 1;
    }
    if (!skip) count_work(d.ctrl, kw);
    const int nwork = c->n_active * spt;
    const bool spec = SPEC && witem < 0 && nwork <= (int)gridDim.x;
    bool more = false;
    const int w = witem >= 0 ? witem : (spec ? (pass == 0 ? wspec : nwork) : xcd_item(nwork, pass, &more));
    if (w >= nwork) skip = 1;
    m.more = more;
    if (!skip) {
        m.q = w / spt;
        m.tile = spec ? tspec : d.active[m.q];
        const int sub = w % spt;
        m.bx0 = (sub % sbx) * B;
        m.by0 = (sub / sbx) * BY;
        m.bxa = min(B, p.nx - m.bx0);
        m.bya = min(BY, p.ny - m.by0);
    }
    m.skip = skip;
    m.dt = dt;
}

// ExcessDeposition (src/Redistribute.f90:158-198) of cell gi, between
// intermed0 = (SH, TA) and intermed3 = (SM[2], TB); also returns Hn_old,
// deltaBt and psiold.
template <class DevT>
__device__ __forceinline__ double excess_deposition(const DevT& d, const Params& p, size_t gi,
                                                    double btc3, double& ho, double& dbt,
                                                    double& psio) {
    const long long FS = d.FS;
    const double go = d.TA.cgam[gi], btc0 = d.TA.btc[gi];
    const double w0 = d.SH[gi], h0 = d.SH[3 * FS + gi];
    ho = compute_hn(w0, d.b0c[gi], btc0, go);
    dbt = btc3 - btc0;
    // intermed0's psi as ComputeDesingularisedVariables left it
    const double hoc = ho < 0.0 ? 0.0 : ho;
    const double hpc = h0 < 0.0 ? 0.0 : h0;
    const double eps = p.hthr;
    const double den = hoc * hoc + dmax(hoc * hoc, eps * eps);
    psio = dmin(2.0 * hoc * hpc / den, p.max_pack);
    const double e = -(h0 * go / (1.0 - p.porosity) - dbt);
    const double lim = -(ho * go - dbt);
    return (e > lim) ? e : lim;
}

// One SSP stage S of MorphodynamicTimeStepper (src/TimeStepper.f90:562-710)
// on one 16 x BY sub-block:
//   E - D of intermed_{S-1} (CalculateMorphodynamicRHS, src/MorphodynamicRHS.f90:72-149)
//   vertex rate (BtSourceTerm, src/MorphodynamicRHS.f90:154-305)
//   bt_S = max(-EroDepth, SSP combination) at this tile's vertices
//   cell topography (ComputeCellCentredTopographicData, :308-360)
//   w, Hnpsi from the change of bt (src/TimeStepper.f90:587-610)
// Stage 3 then runs the checks of src/TimeStepper.f90:713-752.
// ONED: a 1D grid (one row of cells; its vertex rates and cell topography
// take the reference's 1D forms, its upper vertex row keeps a zero rate).
template <int S, int BY, bool ONED, class DevT>
__device__ __forceinline__ bool morpho_pass(const DevT& d, const Params& p, int pass, int witem = -1) {
    constexpr int NT = morpho_nt<BY>();
    constexpr int NH = RW * (BY + 4), NE = MR1 * (BY + 2), NVB = MV * (BY + 1);
    __shared__ double hn[NH];              // Hn of intermed_{S-1}, 2-cell halo
    __shared__ double emd[NE];             // E - D, 1-cell halo
    __shared__ double sbx[NE], sby[NE];
    __shared__ double vb[NVB];             // new vertex bt of this tile's copy
    __shared__ MWork m;
    // a fresh thread index per pass (nothing derived from it is held across
    // passes); the asm has no side effects
    unsigned tidu = threadIdx.x;
    asm("" : "+v"(tidu) : "s"(pass));
    const int tid = (int)(tidu & (NT - 1));
    if (tid == 0) morpho_work<BY>(d, p, m, false, KW_MORPHO + S - 1, pass, witem);
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(m.skip)) return false;   // wave-uniform (a uniform loop)
    const bool more = __builtin_amdgcn_readfirstlane(m.more) != 0;
    const long long FS = d.FS;
    const int t = m.tile, bx0 = m.bx0, by0 = m.by0, bxa = m.bxa, bya = m.bya;
    const double dtm = 2.0 * m.dt;
    const double* Win = (S == 1) ? d.SH : d.SM[S - 2];
    const double* Hin = (S == 1) ? d.SH + 3 * FS : d.SM[S - 2] + FS;
    const auto& Tin = (S == 1) ? d.TA : d.TM[S - 2];
    const auto& Tout = (S == 3) ? d.TB : d.TM[S - 1];
    double* Wout = d.SM[S - 1];
    double* Hout = d.SM[S - 1] + FS;
    const double eps = p.hthr;

    // ---- Hn of the stage state (ComputeDesingularisedVariables) -----------
    for (int q = tid; q < NH; q += NT) {
        const int rx = q % RW - 2, ry = q / RW - 2;
        if (rx > bxa + 1 || ry > bya + 1) continue;
        if (ONED && ry != 0) continue;   // one row of cells
        int tc, ci, cj;
        map_cell(p, t, bx0 + rx, by0 + ry, tc, ci, cj);
        const size_t gi = (size_t)tc * p.NC + cj * p.nx + ci;
        // ghost tiles keep the derived fields they were given: zero
        // (SetDefaultTileData), or the Dirichlet boundary depth.  The inputs
        // are loaded whatever the tile's state (every tile holds the cell)
        // and the value selected, so that the loads do not wait behind the
        // state's branch.
        const bool act = d.tstate[tc] == KST_TILE_ACTIVE;
        double h = compute_hn(Win[gi], d.b0c[gi], Tin.btc[gi], Tin.cgam[gi]);
        if (!act) h = (p.bc == KST_BC_DIRICHLET && dirichlet_ghost(p, tc)) ? p.bc_hn : 0.0;
        hn[q] = h < 0.0 ? 0.0 : h;
    }
    __syncthreads();

    // ---- E - D with the dryness check (src/MorphodynamicRHS.f90:90-145) ---
    for (int q = tid; q < NE; q += NT) {
        const int rx = q % MR1 - 1, ry = q / MR1 - 1;
        if (rx > bxa || ry > bya) continue;
        if (ONED && ry != 0) continue;
        int tc, ci, cj;
        map_cell(p, t, bx0 + rx, by0 + ry, tc, ci, cj);
        const size_t gi = (size_t)tc * p.NC + cj * p.nx + ci;
        // the cell's inputs, loaded with its state (every tile holds the
        // cell) and held by the empty asm, so that the loads issue together
        // rather than one after another behind the state's branch
        const double bxc = Tin.cbx[gi], byc = Tin.cby[gi];
        const double hin = Hin[gi], btc = Tin.btc[gi];
        const double ua = (S == 1) ? d.S3[FS + gi] : d.UV[gi];
        const double va = (S == 1) ? (ONED ? 0.0 : d.S3[2 * FS + gi]) : d.UV[FS + gi];
        const int st = d.tstate[tc];
        asm volatile("" ::"v"(st), "v"(bxc), "v"(byc), "v"(hin), "v"(btc), "v"(ua), "v"(va));
        sbx[q] = bxc;
        sby[q] = byc;
        double e = 0.0;
        const bool on = st == KST_TILE_ACTIVE;
        const bool own = rx >= 0 && ry >= 0 && rx < bxa && ry < bya;
        if (on) {
            const double h = hn[rg(rx, ry)];
            double hp = hin;
            if (hp < 0.0) hp = 0.0;
            const double den = h * h + dmax(h * h, eps * eps);
            const double psi = dmin(2.0 * h * hp / den, p.max_pack);
            // the velocities intermed_k holds: desingularised from the first
            // operator's stage-3 momentum, never recomputed in M
            double u, v;
            if (S == 1) {
                const double rho = density(p, psi);
                u = 2.0 * h * ua / (h * h + dmax(h * h, eps * eps)) / rho;
                v = ONED ? 0.0 : 2.0 * h * va / (h * h + dmax(h * h, eps * eps)) / rho;
                if (own) { d.UV[gi] = u; d.UV[FS + gi] = v; }
            } else {
                u = ua;
                v = va;
            }
            // the S and N neighbours only on a 2D grid (src/MorphodynamicRHS.f90:113-135)
            if (!(h < eps || hn[rg(rx - 1, ry)] < eps || hn[rg(rx + 1, ry)] < eps ||
                  (!ONED && (hn[rg(rx, ry - 1)] < eps || hn[rg(rx, ry + 1)] < eps)))) {
                CellState cs{h, u, v, psi, btc, bxc, byc};
                e = erosion_minus_deposition(p, cs);
            }
        }
        emd[q] = e;
    }
    __syncthreads();

    // ---- vertex rates and the stage update of this tile's copy ------------
    const int NV1 = p.nx + 1;
    const size_t NV = (size_t)(p.nx + 1) * (p.ny + 1);
    const double psib = 1.0 - p.porosity;
    for (int q = tid; q < NVB; q += NT) {
        const int a = q % MV, b = q / MV;
        if (a > bxa || b > bya) continue;
        const int vi = bx0 + a, vj = by0 + b;
        double rate;
        if constexpr (ONED) {
            // BtSourceTerm on a 1D grid (src/MorphodynamicRHS.f90:272-302):
            // cells vi-1 and vi of the row; at a tile edge without an active
            // neighbour the cell's own half; the upper vertex row keeps a zero rate
            rate = 0.0;
            if (b == 0) {
                const int cl = MR1 + a, cr = MR1 + a + 1;   // cells vi-1, vi (row 0)
                bool both = true;
                if (vi == 0) both = d.tstate[nb_tile(p, t, -1, 0)] == KST_TILE_ACTIVE;
                if (vi == p.nx) both = d.tstate[nb_tile(p, t, 1, 0)] == KST_TILE_ACTIVE;
                if (both) {
                    const double bx = 0.5 * (sbx[cl] + sbx[cr]);
                    const double gam = gamma_of(p.geometric, bx, 0.0);
                    rate = -0.5 * gam * (emd[cl] + emd[cr]) / psib;
                } else {
                    const int cc = (vi == 0) ? cr : cl;   // the tile's own edge cell
                    const double bx = 0.5 * sbx[cc];
                    const double gam = gamma_of(p.geometric, bx, 0.0);
                    rate = -0.5 * gam * emd[cc] / psib;
                }
            }
        } else {
        // cells (vi-1, vj-1), (vi-1, vj), (vi, vj-1), (vi, vj)
        const int i00 = b * MR1 + a, i01 = (b + 1) * MR1 + a;
        const int i10 = b * MR1 + a + 1, i11 = (b + 1) * MR1 + a + 1;
        Kahan kx, ky, ke;
        kx.add(sbx[i00]); kx.add(sbx[i01]);
        if (vi == p.nx && vj == p.ny) kx.add(sbx[i10] + sbx[i11]);   // src/MorphodynamicRHS.f90:266
        else { kx.add(sbx[i10]); kx.add(sbx[i11]); }
        ky.add(sby[i00]); ky.add(sby[i01]); ky.add(sby[i10]); ky.add(sby[i11]);
        ke.add(emd[i00]); ke.add(emd[i01]);
        if (vi == 0 && vj >= 1 && vj <= p.ny - 1) ke.add(emd[i10] + emd[i11]);   // :210
        else { ke.add(emd[i10]); ke.add(emd[i11]); }
        const double gam = gamma_of(p.geometric, 0.25 * kx.s, 0.25 * ky.s);
        rate = -0.25 * gam / psib * ke.s;
        }
        const size_t vidx = (size_t)t * NV + vj * NV1 + vi;
        const double bt0 = d.TA.vbt[vidx];
        double bn;
        if (S == 1) bn = bt0 + dtm * rate;
        else if (S == 2) bn = 0.75 * bt0 + 0.25 * (Tin.vbt[vidx] + dtm * rate);
        else bn = (1.0 / 3.0) * bt0 + (2.0 / 3.0) * (Tin.vbt[vidx] + dtm * rate);
        bn = dmax(-p.ero_depth, bn);
        vb[q] = bn;
        if ((a < bxa || vi == p.nx) && (b < bya || vj == p.ny)) Tout.vbt[vidx] = bn;
    }
    __syncthreads();

    // ---- cell topography, w and Hnpsi of the owned cells -------------------
    const int tx = tid % B, ty = tid / B;
    if (tx >= bxa || ty >= bya) return more;
    const int i = bx0 + tx, j = by0 + ty;
    const double* b0 = d.vb0 + (size_t)t * NV;
#define VB(a, b) vb[(b) * MV + (a)]
#define V0(ii, jj) b0[(jj) * NV1 + (ii)]
    const double t00 = VB(tx, ty), t10 = VB(tx + 1, ty), t01 = VB(tx, ty + 1), t11 = VB(tx + 1, ty + 1);
    double btc, dbdx, dbdy;
    if constexpr (ONED) {
        // src/MorphodynamicRHS.f90:348-357
        btc = 0.5 * (t00 + t10);
        Kahan kx;
        kx.add(V0(i + 1, j)); kx.add(t10); kx.add(-V0(i, j)); kx.add(-t00);
        dbdx = p.dxr * kx.s;
        dbdy = 0.0;
    } else {
        Kahan kt;
        kt.add(t00); kt.add(t10); kt.add(t01); kt.add(t11);
        btc = 0.25 * kt.s;
        Kahan kx;
        kx.add(V0(i + 1, j)); kx.add(t10); kx.add(-V0(i, j)); kx.add(-t00);
        kx.add(V0(i + 1, j + 1)); kx.add(t11); kx.add(-V0(i, j + 1)); kx.add(-t01);
        dbdx = 0.5 * p.dxr * kx.s;
        Kahan ky;
        ky.add(V0(i, j + 1)); ky.add(t01); ky.add(-V0(i, j)); ky.add(-t00);
        ky.add(V0(i + 1, j + 1)); ky.add(t11); ky.add(-V0(i + 1, j)); ky.add(-t10);
        dbdy = 0.5 * p.dyr * ky.s;
    }
#undef VB
#undef V0
    const double gn = gamma_of(p.geometric, dbdx, dbdy);
    const size_t gi = (size_t)t * p.NC + j * p.nx + i;
    const double b0c = d.b0c[gi];
    const double w0 = d.SH[gi], h0 = d.SH[3 * FS + gi];
    const double btc0 = d.TA.btc[gi], go = d.TA.cgam[gi];
    const double ho = compute_hn(w0, b0c, btc0, go);
    const double db = btc - btc0;
    double w = -db / gn / gn;
    w = w + btc;
    w = w + ho * go / gn / gn;
    w = w + b0c;
    double hp = -(1.0 - p.porosity) * db / gn;
    hp = hp + h0 * go / gn;
    Wout[gi] = w;
    Hout[gi] = hp;
    Tout.btc[gi] = btc;
    Tout.cbx[gi] = dbdx;
    Tout.cby[gi] = dbdy;
    Tout.cgam[gi] = gn;

    if (S == 3) {
        // excess deposition and depth-change checks (src/TimeStepper.f90:713-752)
        const double hnew = compute_hn(w, b0c, btc, gn);
        double ho2, dbt, psio;
        const double ex = excess_deposition(d, p, gi, btc, ho2, dbt, psio);
        Ctrl* c = d.ctrl;
        if (ex > kEps && dbt > kEps && psio > -kEps) {
            if (ho2 < p.ero_hcrit || fabs(hnew - ho2) < p.ero_hcrit) {
                const int k = atomicAdd(&c->ncand, 1);
                if (k < d.cand_cap) {
                    d.cand_key[k] = (long long)m.q * p.NC + j * p.nx + i;
                    d.cand_ex[k] = ex;
                } else {
                    atomicOr(&c->mref, 4);
                }
            } else {
                atomicOr(&c->mref, 1);
            }
        } else if (!(ho2 < p.ero_hcrit)) {
            const double rel = fabs(hnew - ho2) / fabs(ho2);
            if (rel > 0.1) atomicOr(&c->mref, 1);
        }
    }
    return more;
}

// The grid is sized for the active set the host last saw; a block whose
// share of the list holds more items takes them in turn.
template <int S, int BY = B, bool ONED = false>
// (the 16 x 16 shape keeps 4 waves per SIMD, two of its 512-thread blocks
// per CU: at most 128 VGPRs)
__global__ __launch_bounds__(morpho_nt<BY>(), BY == B ? 4 : 1) void morpho_kernel(const Dev* __restrict__ dp,
                                                               const Params* __restrict__ pp) {
    for (int pass = 0;; ++pass) {
        // fresh argument pointers per pass: loads through them are not
        // hoisted out of the loop and held in registers across passes
        // (Params through the constant address space: through a plain
        // pointer from the asm its fields would be flat vector loads)
        const Dev* d = dp;
        ConstArg<Params>* p = (ConstArg<Params>*)pp;
        asm("" : "+s"(d), "+s"(p) : "s"(pass));
        if (!morpho_pass<S, BY, ONED>(const_arg(d), generic(*p), pass)) break;
        __syncthreads();
    }
}

// A write into a ghost tile's part of intermed3, listed for the next step's
// refresh.  Only the redistribution workgroup writes the list: it counts the
// entries in LDS (nd, loaded from and stored back to Ctrl::ndirty by the
// kernel), since a returning device-scope atomic per entry would put a
// device round trip on the critical path of every phase that writes one.
template <class DevT>
__device__ __forceinline__ void note_dirty(const DevT& d, long long e, unsigned long long* nd) {
    const unsigned long long q = atomicAdd(nd, 1ULL);
    if (q < (unsigned long long)d.dirty_cap) d.dirty[q] = e;
}

// The 3x3 update of RedistributeCell (src/Redistribute.f90:402-470) for one
// cell: topography from its four vertex bt (t) and b0 (e), then w and Hnpsi
// from intermed0's Hn, Hnpsi and gamma (cbt0, cgo, cw0, ch0, cb0; in ghost
// tiles, whose derived fields the reference never computes, Hn is hghost:
// zero, or the Dirichlet boundary depth).
struct CellRe {
    double btc, dbdx, dbdy, gn, w, hp;
};

__device__ __forceinline__ CellRe recompute_cell(const Params& p, double t00, double t10, double t01,
                                                 double t11, double e00, double e10, double e01,
                                                 double e11, double cbt0, double cgo, double cw0,
                                                 double ch0, double cb0, bool cact,
                                                 double hghost = 0.0) {
    if (p.oned) {
        // a 1D grid: the lower vertex row and the 1D order of the w update
        // (src/MorphodynamicRHS.f90:408-414, src/Redistribute.f90:444-470)
        CellRe r;
        r.btc = 0.5 * (t00 + t10);
        Kahan kx;
        kx.add(e10); kx.add(t10); kx.add(-e00); kx.add(-t00);
        r.dbdx = p.dxr * kx.s;
        r.dbdy = 0.0;
        r.gn = gamma_of(p.geometric, r.dbdx, 0.0);
        const double dbb = r.btc - cbt0;
        double ho = hghost;
        if (cact) {
            ho = compute_hn(cw0, cb0, cbt0, cgo);
            if (ho < 0.0) ho = 0.0;
        }
        double w = -dbb / r.gn / r.gn;
        w = w + ho * cgo / r.gn / r.gn;
        w = w + r.btc;
        r.w = w + cb0;
        r.hp = ch0 * cgo / r.gn - (1.0 - p.porosity) * dbb / r.gn;
        return r;
    }
    Kahan kt;
    kt.add(t00); kt.add(t10); kt.add(t01); kt.add(t11);
    Kahan kx;
    kx.add(e10); kx.add(t10); kx.add(-e00); kx.add(-t00);
    kx.add(e11); kx.add(t11); kx.add(-e01); kx.add(-t01);
    Kahan ky;
    ky.add(e01); ky.add(t01); ky.add(-e00); ky.add(-t00);
    ky.add(e11); ky.add(t11); ky.add(-e10); ky.add(-t10);
    CellRe r;
    r.btc = 0.25 * kt.s;
    r.dbdx = 0.5 * p.dxr * kx.s;
    r.dbdy = 0.5 * p.dyr * ky.s;
    r.gn = gamma_of(p.geometric, r.dbdx, r.dbdy);
    const double dbb = r.btc - cbt0;
    double ho = hghost;
    if (cact) {
        ho = compute_hn(cw0, cb0, cbt0, cgo);
        if (ho < 0.0) ho = 0.0;
    }
    double w = r.btc;
    w = w + (ho * cgo / r.gn - dbb / r.gn) / r.gn;
    r.w = w + cb0;
    r.hp = ch0 * cgo / r.gn - (1.0 - p.porosity) * dbb / r.gn;
    return r;
}

// One corrected vertex (di, dj) of tile k: the tile's own copy and those of
// the tiles that share it (src/Redistribute.f90:330-398).  Copies in ghost
// tiles are listed for the next step's refresh and counted.
// With write = false the copies are only counted.
template <class DevT>
__device__ __forceinline__ void put_vertex(const DevT& d, const Params& p, int k, int di, int dj,
                                           double val, unsigned long long* ghost,
                                           unsigned long long* nd, bool write = true) {
    const int nx = p.nx, ny = p.ny, NV1 = nx + 1;
    const size_t NV = (size_t)(nx + 1) * (ny + 1);
    double* bt3 = d.TB.vbt;
    auto put = [&](int tile, int ii, int jj) {
        const size_t v = (size_t)tile * NV + jj * NV1 + ii;
        if (write) bt3[v] = val;
        if (d.tstate[tile] != KST_TILE_ACTIVE) {
            if (write) note_dirty(d, (long long)v | (1LL << 62), nd);
            ++*ghost;
        }
    };
    if (write) bt3[(size_t)k * NV + dj * NV1 + di] = val;
    if (di == 0) put(nb_tile(p, k, -1, 0), nx, dj);
    else if (di == nx) put(nb_tile(p, k, 1, 0), 0, dj);
    if (p.oned) return;   // no row copies on a 1D grid (src/Redistribute.f90:369)
    if (dj == 0) {
        put(nb_tile(p, k, 0, -1), di, ny);
        if (di == 0) put(nb_tile(p, k, -1, -1), nx, ny);
        if (di == nx) put(nb_tile(p, k, 1, -1), 0, ny);
    }
    if (dj == ny) {
        put(nb_tile(p, k, 0, 1), di, 0);
        if (di == 0) put(nb_tile(p, k, -1, 1), nx, 0);
        if (di == nx) put(nb_tile(p, k, 1, 1), 0, 0);
    }
}

// The excess deposition of a candidate re-evaluated against the current bed
// and the vertex correction that removes it (src/Redistribute.f90:158-198,
// 280-330): returns 0 if no correction is needed, -1 if the cell borders no
// depositional vertex, else 1 with the new values nv of the N depositional
// vertices (di, dj) in vertex order.  x3/x0: the cell's vertex bt now and in
// intermed0; btc3/w3: its bt and w now.  *delta_out, when given, receives
// the factor: vertex k with b_k = x3[k] - x0[k] > 0 moves to x3[k] - delta b_k.
__device__ __forceinline__ int vertex_correction(const Params& p, int i, int j, const double* x3,
                                                 const double* x0, double gamold, double btc0,
                                                 double btc3, double w0, double h0, double b0c,
                                                 double w3, double* nv, int* di, int* dj, int& N,
                                                 double* delta_out = nullptr) {
    const int vi[4] = {i, i + 1, i, i + 1}, vj[4] = {j, j, j + 1, j + 1};
    const double ho_raw = compute_hn(w0, b0c, btc0, gamold);
    const double db = btc3 - btc0;
    double corr;
    {
        const double e = -(h0 * gamold / (1.0 - p.porosity) - db);
        const double lim = -(ho_raw * gamold - db);
        corr = (e > lim) ? e : lim;
    }
    N = 0;
    if (!(corr > kEps)) return 0;
    double bd[4];
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double bq = x3[q] - x0[q];
        bd[q] = 0.0; di[q] = -10; dj[q] = -10; nv[q] = 0.0;
        if (bq > 0.0) {
            // compact the depositional vertices in order (N is at most q)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (r == N) { bd[r] = bq; di[r] = vi[q]; dj[r] = vj[q]; nv[r] = x3[q]; }
            sum = sum + bq;
            N++;
        }
    }
    if (N == 0) return -1;
    double delta = 4.0 * corr / sum;
    const double hnold = ho_raw < 0.0 ? 0.0 : ho_raw;   // intermed0's stored Hn
    double tol = kEps * w3 * 10.0;
    double adj = 0.0;
    {
        Kahan ks;
        ks.add(hnold * gamold); ks.add(-db); ks.add(corr);
        if (fabs(ks.s) < tol) {
            Kahan ka;
            ka.add(tol); ka.add(-hnold * gamold); ka.add(db);
            adj = ka.s;
            adj = adj * (4.0 / sum);
            adj = adj - delta;
            adj = dmax(adj, 0.0);
        }
    }
    const double hgpo = h0 * gamold / (1.0 - p.porosity);
    tol = kEps * 10.0;
    {
        Kahan ks;
        ks.add(hgpo); ks.add(-db); ks.add(corr);
        if (fabs(ks.s) < tol) {
            Kahan ka;
            ka.add(tol); ka.add(-hgpo); ka.add(db);
            double ap = ka.s;
            ap = ap * (4.0 / sum);
            ap = ap - delta;
            adj = dmax(adj, ap);
        }
    }
    delta = delta + adj;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (q < N) nv[q] = nv[q] - delta * bd[q];
    if (delta_out) *delta_out = delta;
    return 1;
}

constexpr int RD_THREADS = 1024;      // candidates walked at once (one per thread)
constexpr int RD_LDS = 2048;          // candidates sorted in LDS (more: in global memory)

// Neighbour k (0..7) of a cell at offset (kDI[k], kDJ[k]).
__constant__ const int kDI[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
__constant__ const int kDJ[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
// Vertex k = a + 2b of a cell (a, b in 0..1) is shared by the neighbours at
// offsets (a - 1 .. a, b - 1 .. b) other than the cell: their indices m in
// the list above, and the index of the vertex among each one's own.
constexpr int kNbM[4][3] = {{0, 1, 3}, {1, 2, 4}, {3, 5, 6}, {4, 6, 7}};
constexpr int kNbV[4][3] = {{3, 2, 1}, {3, 2, 0}, {3, 1, 0}, {2, 1, 0}};

// The walk over candidates base .. base+n-1 (n <= RD_THREADS) of the list in
// cand_key/cand_ex, which need not be sorted: they are ranked here.  Returns
// true if it stopped at a candidate that borders no depositional vertex.
template <class DevT>
__device__ bool redistribute_small(const DevT& d, const Params& p, int base, int n, double* lx,
                                   long long* lk, int* lev, int* cellid, unsigned long long* nd) {
    __shared__ double slot[RD_THREADS][4];     // vertex bt of each candidate's cell
    __shared__ unsigned char rec[RD_THREADS];  // its bt and w were recomputed by a correction
    __shared__ unsigned char res[RD_THREADS];  // 0 none, 1 corrected, 2 no depositional vertex
    __shared__ unsigned char dmask[RD_THREADS];
    __shared__ int s_stop;
    __shared__ unsigned long long s_done, s_ghost;
    Ctrl* c = d.ctrl;
    const int tid = threadIdx.x;
    const long long FS = d.FS;
    const int nx = p.nx, ny = p.ny, NV1 = nx + 1;
    const int NX = p.ntx * nx, NY = p.nty * ny;
    const size_t NV = (size_t)(nx + 1) * (ny + 1);
    if (tid == 0) { s_stop = n; s_done = 0; s_ghost = 0; }
    // each candidate's tile is read from the active list here and used after
    // the ranking, which that load overlaps
    int ctile = 0, ccell = 0;
    if (tid < n) {
        const long long kk = d.cand_key[base + tid];
        lk[tid] = kk; lx[tid] = d.cand_ex[base + tid]; lev[tid] = 0;
        ctile = d.active[(int)(kk / p.NC)];
        ccell = (int)(kk % p.NC);
    }
    __syncthreads();
    // ---- rank of each candidate in (excess, scan order): P threads per
    // candidate count over interleaved parts of the list
    {
        int P = 1;
        while (P < 16 && 2 * P * n <= RD_THREADS) P <<= 1;
        const int q = tid % n, part = tid / n;
        if (n > 0 && part < P) {
            const double x = lx[q];
            const long long kk = lk[q];
            int r = 0;
#pragma unroll 8
            for (int u = part; u < n; u += P) {
                const double xu = lx[u];
                const long long ku = lk[u];
                r += (int)((xu < x) | ((xu == x) & (ku < kk)));
            }
            atomicAdd(&lev[q], r);
        }
        __syncthreads();
        if (tid < n) cellid[lev[tid]] = ctile * p.NC + ccell;
    }
    __syncthreads();
    int k = 0, i = 0, j = 0, I = 0, J = 0;
    if (tid < n) {
        const int g = cellid[tid];
        k = g / p.NC;
        i = (g % p.NC) % nx;
        j = (g % p.NC) / nx;
        I = (k % p.ntx) * nx + i;
        J = (k / p.ntx) * ny + j;
        d.cand_pos[g] = tid;
    }
    __syncthreads();
    // ---- neighbours within one cell, and the per-candidate state
    double x0[4], e[4];
    double gamold = 1.0, btc0 = 0.0, w0 = 0.0, h0 = 0.0, b0c = 0.0, btc3s = 0.0, w3s = 0.0;
    int nb[8];                                 // this thread's neighbours, kept in registers
#pragma unroll
    for (int q = 0; q < 8; ++q) nb[q] = -1;
    if (tid < n) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            int I2 = I + kDI[q], J2 = J + kDJ[q];
            if (p.periodic) {   // the neighbours across the domain edge (along x only on a 1D grid)
                I2 += I2 < 0 ? NX : (I2 >= NX ? -NX : 0);
                if (!p.oned) J2 += J2 < 0 ? NY : (J2 >= NY ? -NY : 0);
            }
            int r = -1;
            if (I2 >= 0 && J2 >= 0 && I2 < NX && J2 < NY)
                r = d.cand_pos[((J2 / ny) * p.ntx + I2 / nx) * p.NC + (J2 % ny) * nx + I2 % nx];
            nb[q] = r;
        }
        const size_t gc = (size_t)k * p.NC + j * nx + i;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const size_t v = (size_t)k * NV + (j + (q >> 1)) * NV1 + i + (q & 1);
            slot[tid][q] = d.TB.vbt[v];
            x0[q] = d.TA.vbt[v];
            e[q] = d.vb0[v];
        }
        gamold = d.TA.cgam[gc]; btc0 = d.TA.btc[gc]; btc3s = d.TB.btc[gc];
        w0 = d.SH[gc]; h0 = d.SH[3 * FS + gc]; b0c = d.b0c[gc]; w3s = d.SM[2][gc];
        rec[tid] = 0;
        res[tid] = 0;
        dmask[tid] = 0;
        lev[tid] = 0;
    }
    __syncthreads();
    // ---- decisions in rounds: a candidate is decided once every earlier
    // neighbour has been; lev records the round (0: not yet)
    for (int round = 1;; ++round) {
        int pending = 0;
        if (tid < n && lev[tid] == 0) {
            bool ready = true;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int r = nb[m];
                if (r >= 0 && r < tid && (lev[r] == 0 || lev[r] == round)) ready = false;
            }
            if (!ready) {
                pending = 1;
            } else {
                double x3[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) x3[q] = slot[tid][q];
                double btc3 = btc3s, w3 = w3s;
                if (rec[tid]) {
                    const CellRe r = recompute_cell(p, x3[0], x3[1], x3[2], x3[3], e[0], e[1], e[2],
                                                    e[3], btc0, gamold, w0, h0, b0c, true);
                    btc3 = r.btc;
                    w3 = r.w;
                }
                double nv[4], delta;
                int di[4], dj[4], N;
                const int out = vertex_correction(p, i, j, x3, x0, gamold, btc0, btc3, w0, h0, b0c,
                                                  w3, nv, di, dj, N, &delta);
                if (out < 0) {
                    res[tid] = 2;
                    atomicMin(&s_stop, tid);
                } else if (out > 0) {
                    res[tid] = 1;
                    // the depositional vertices k (in vertex order, as
                    // vertex_correction compacts them) move; the later
                    // candidates around each take the new value into their
                    // copy: vertex k of this cell is vertex kNbV[k][u] of
                    // neighbour kNbM[k][u]
                    int mask = 0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const double bq = x3[k] - x0[k];
                        if (bq > 0.0) {
                            const double v = x3[k] - delta * bq;
                            mask |= 1 << k;
                            slot[tid][k] = v;
#pragma unroll
                            for (int u = 0; u < 3; ++u) {
                                const int r = nb[kNbM[k][u]];
                                if (r > tid) slot[r][kNbV[k][u]] = v;
                            }
                        }
                    }
                    // the 3x3 update recomputes every later neighbour's bt and w
#pragma unroll
                    for (int m = 0; m < 8; ++m)
                        if (nb[m] > tid) rec[nb[m]] = 1;
                    dmask[tid] = (unsigned char)mask;
                }
                lev[tid] = round;
            }
        }
        if (__syncthreads_count(pending) == 0) break;
    }
    const int stop = s_stop;
    // ---- vertex copies of the corrections before the stop: each vertex is
    // written by the last correction that moves it; every correction's
    // copies in ghost tiles are counted
    unsigned long long ghost = 0;
    if (tid < stop && res[tid] == 1) {
        const int mask = dmask[tid];
        for (int q = 0; q < 4; ++q) {
            if (!(mask & (1 << q))) continue;
            const int a = q & 1, b = q >> 1;
            bool last = true;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const int r = nb[m];
                if (r <= tid || r >= stop || res[r] != 1) continue;
                const int a2 = a - kDI[m], b2 = b - kDJ[m];
                if (a2 >= 0 && a2 <= 1 && b2 >= 0 && b2 <= 1 && (dmask[r] >> (a2 + 2 * b2)) & 1) last = false;
            }
            put_vertex(d, p, k, i + a, j + b, slot[tid][q], &ghost, nd, last);
        }
    }
    __syncthreads();
    // ---- the 3x3 cells of those corrections from the final vertices
    for (int u = tid; u < 9 * stop; u += RD_THREADS) {
        const int q = u / 9, lane = u % 9;
        if (res[q] != 1) continue;
        if (p.oned && lane % 3 != 1) continue;   // the row's three cells on a 1D grid
        const int g = cellid[q];
        const int kq = g / p.NC, iq = (g % p.NC) % nx, jq = (g % p.NC) / nx;
        int ct, ci, cj;
        map_cell(p, kq, iq + lane / 3 - 1, jq + lane % 3 - 1, ct, ci, cj);
        const size_t gcl = (size_t)ct * p.NC + cj * nx + ci;
        const double* vb = d.vb0 + (size_t)ct * NV;
        const double* vt = d.TB.vbt + (size_t)ct * NV;
        const bool cact = d.tstate[ct] == KST_TILE_ACTIVE;
        const CellRe r = recompute_cell(p, vt[cj * NV1 + ci], vt[cj * NV1 + ci + 1], vt[(cj + 1) * NV1 + ci],
                                        vt[(cj + 1) * NV1 + ci + 1], vb[cj * NV1 + ci], vb[cj * NV1 + ci + 1],
                                        vb[(cj + 1) * NV1 + ci], vb[(cj + 1) * NV1 + ci + 1], d.TA.btc[gcl],
                                        d.TA.cgam[gcl], d.SH[gcl], d.SH[3 * FS + gcl], d.b0c[gcl], cact,
                                        (!cact && p.bc == KST_BC_DIRICHLET && dirichlet_ghost(p, ct)) ? p.bc_hn : 0.0);
        if (cact) {
            d.TB.btc[gcl] = r.btc; d.TB.cbx[gcl] = r.dbdx; d.TB.cby[gcl] = r.dbdy; d.TB.cgam[gcl] = r.gn;
            d.SM[2][gcl] = r.w;
            d.SM[2][FS + gcl] = r.hp;
        } else {
            d.S3[gcl] = r.w;
            d.S3[3 * FS + gcl] = r.hp;
            note_dirty(d, (long long)gcl, nd);
            ++ghost;
        }
    }
    if (tid < n) d.cand_pos[cellid[tid]] = -1;
    if (tid < stop && res[tid] == 1) atomicAdd(&s_done, 1ULL);
    if (ghost) atomicAdd(&s_ghost, ghost);
    __syncthreads();
    if (tid == 0) {
        c->redist_cells += (long long)s_done;
        c->redist_ghost += (long long)s_ghost;
        if (stop < n) c->mref |= 2;
    }
    return stop < n;
}

// AddToRedistList + RedistributeGrid (src/Redistribute.f90:69-244): the
// candidates the checks marked, sorted by excess deposition (ties in the
// check loop's order: active-list position, row, column), each re-evaluated
// and corrected in turn.  One workgroup.
//
// Up to RD_THREADS candidates (the common case) are handled with the bed
// state in LDS.  A correction's decision reads only the cell's four vertex bt
// and its bt and w, which are the stored stage values until a correction of
// a cell within one cell (a shared vertex) recomputes them from those
// vertices.  So thread q owns sorted candidate q: the candidates are decided
// in rounds (a candidate once every earlier candidate within one cell is)
// from the vertex values held in LDS (a correction updates the copies held
// by the later candidates that share the vertex), and every decision sees
// the state the sequential walk gives it.  The walk stops at the first
// candidate that borders no depositional vertex (the attempt is then
// refined); only the corrections before it are written: the vertex copies,
// then the 3x3 cells around each correction from the final vertices, which
// is the value the walk leaves in them (a later correction that moves one of
// a cell's vertices recomputes that cell).
//
// Larger sets are sorted first (bitonic, in LDS up to RD_LDS candidates, else
// in place in global memory) and walked in consecutive chunks of the sorted
// list, each chunk as above on the state the chunks before it wrote, until a
// chunk stops: the same walk, in the same order, as the reference's.
template <class DevT>
__device__ __forceinline__ void redistribute_body(const DevT& d, const Params& p) {
    __shared__ int sh_n, sh_skip;
    __shared__ unsigned long long sh_ndirty;
    __shared__ double lx[RD_LDS];
    __shared__ long long lk[RD_LDS];
    __shared__ int lev[RD_THREADS];
    __shared__ int list[RD_THREADS];
    Ctrl* c = d.ctrl;
    const int tid = threadIdx.x;
    if (tid == 0) {
        const CtrlSnap cs = snap_ctrl(ctrl_view(d));
        const long long ndirty = ctrl_view(d)->ndirty;
        int skip = (cs.status != ST_NEW && cs.status != ST_RETRY);
        if (!skip) {
            const double dt = step_dt(d, p, &cs);
            if (!first_operator_ok(&cs, p, dt) || cs.mref) skip = 1;
        }
        if (!skip && cs.ncand <= 0) skip = 1;
        if (!skip) count_work(c, KW_REDIST);
        sh_skip = skip;
        sh_n = cs.ncand;
        sh_ndirty = ndirty;
    }
    __syncthreads();
    if (sh_skip) return;
    const int n = min(sh_n, d.cand_cap);
    const int chunk = d.rd_chunk;
    if (n > chunk) {
        // ---- sort by (excess, scan order); the padding sorts last
        int m2 = 1;
        while (m2 < n) m2 <<= 1;
        long long* K = d.cand_key;
        double* X = d.cand_ex;
        const bool in_lds = m2 <= d.rd_lds_sort;
        if (in_lds) {
            for (int q = tid; q < m2; q += RD_THREADS) {
                lk[q] = q < n ? K[q] : 0x7fffffffffffffffLL;
                lx[q] = q < n ? X[q] : kHuge;
            }
        } else {
            for (int q = n + tid; q < m2; q += RD_THREADS) {
                K[q] = 0x7fffffffffffffffLL;
                X[q] = kHuge;
            }
        }
        __syncthreads();
        long long* SK = in_lds ? lk : K;
        double* SX = in_lds ? lx : X;
        for (int kk = 2; kk <= m2; kk <<= 1) {
            for (int jj = kk >> 1; jj > 0; jj >>= 1) {
                for (int q = tid; q < m2; q += RD_THREADS) {
                    const int l = q ^ jj;
                    if (l <= q) continue;
                    const double xq = SX[q], xl = SX[l];
                    const long long kq = SK[q], kl = SK[l];
                    const bool l_less = (xl < xq) || (xl == xq && kl < kq);
                    const bool up = (q & kk) == 0;
                    if (l_less == up) {
                        SX[q] = xl; SX[l] = xq;
                        SK[q] = kl; SK[l] = kq;
                    }
                }
                __syncthreads();
            }
        }
        if (in_lds) {
            for (int q = tid; q < n; q += RD_THREADS) {
                K[q] = lk[q];
                X[q] = lx[q];
            }
        }
        __syncthreads();
    }
    // ---- the walk, chunk by chunk (a chunk reads what the earlier ones wrote:
    // one workgroup, so the barrier orders its global stores and loads); one
    // call site, so that the walk is inlined with its arrays in registers
    for (int b = 0;; b += chunk) {
        const bool stopped = redistribute_small(d, p, b, min(chunk, n - b), lx, lk, lev, list, &sh_ndirty);
        __syncthreads();
        if (stopped || b + chunk >= n) break;
    }
    if (tid == 0) c->ndirty = (long long)sh_ndirty;
}

__global__ __launch_bounds__(RD_THREADS) void redistribute_kernel(const Dev* __restrict__ dp, const Params* __restrict__ pp) {
    redistribute_body(const_arg(dp), *pp);
}

// ComputeInterfacialTopographicData on the new bed (src/TimeStepper.f90:774-779),
// plus its CFL numerators and prefactors, in (tile, part) items.

// Whether the face topography of the new bed is due in this attempt (the
// morphodynamic operator ran and its checks passed).
template <class DevT>
__device__ __forceinline__ bool face_topo_due(const DevT& d, const Params& p) {
    const CtrlSnap cs = snap_ctrl(ctrl_view(d));
    const CtrlSnap* c = &cs;
    if (c->status != ST_NEW && c->status != ST_RETRY) return false;
    const double dt = step_dt(d, p, c);
    return first_operator_ok(c, p, dt) && !c->mref;
}

// (tile, part) item `it` of the face topography of the new bed.
template <class DevT>
__device__ __forceinline__ void face_topo_item(const DevT& d, const Params& p, int it) {
    const int t = d.active[it / TILE_PARTS], part = it % TILE_PARTS;
    // a face's static terms read only that face's topography, computed in
    // the same pass, and cell gamma, which the morphodynamic stage wrote
    tile_face_topo<true>(d, p, d.TB, t, part, TILE_PARTS);
    tile_prefactors(d, p, d.TB, t, part, TILE_PARTS);
}

__global__ __launch_bounds__(256) void face_topo_kernel(const Dev* __restrict__ dp, const Params* __restrict__ pp) {
    const auto& d = const_arg(dp);
    const Params& p = *pp;
    __shared__ int sh_skip, sh_n;
    const auto* c = ctrl_view(d);
    if (threadIdx.x == 0) {
        const int skip = face_topo_due(d, p) ? 0 : 1;
        if (!skip) count_work(d.ctrl, KW_FACETOPO);
        sh_skip = skip;
        sh_n = c->n_active * TILE_PARTS;
    }
    __syncthreads();
    if (sh_skip) return;
    // (tile, part) items, the grid sized for the active set the host last saw
    for (int it = blockIdx.x; it < sh_n; it += gridDim.x) face_topo_item(d, p, it);
}

// One of the TILE_PARTS shares of tile t's bed copy.
template <class DevT>
__device__ void commit_part(const DevT& d, const Params& p, int t, int part) {
    const int lane0 = threadIdx.x + part * blockDim.x, stride = blockDim.x * TILE_PARTS;
    const size_t NV = (size_t)(p.nx + 1) * (p.ny + 1);
    const int nv = (int)NV, nmax = max(max(nv, p.NC), max(p.NFX, p.NFY));
    // every array's element k read first, then written: written group by
    // group, each group's loads would wait for the stores before them (the
    // two generations are not known to be distinct)
    for (int k = lane0; k < nmax; k += stride) {
        const bool hv = k < nv, hc = k < p.NC, hx = k < p.NFX, hy = k < p.NFY;
        const size_t v = (size_t)t * NV + k, g = (size_t)t * p.NC + k;
        const size_t fx = (size_t)t * p.NFX + k, fy = (size_t)t * p.NFY + k;
        double a = 0.0, c[7] = {}, x[6] = {}, y[6] = {};
        if (hv) a = d.TB.vbt[v];
        if (hc) {
            c[0] = d.TB.btc[g]; c[1] = d.TB.cbx[g]; c[2] = d.TB.cby[g]; c[3] = d.TB.cgam[g];
            c[4] = d.TB.pxu[g]; c[5] = d.TB.pxv[g]; c[6] = d.TB.pyv[g];
        }
        if (hx) {
            x[0] = d.TB.xbt[fx]; x[1] = d.TB.xby[fx]; x[2] = d.TB.xgam[fx];
            x[3] = d.TB.xib[fx]; x[4] = d.TB.xnp[fx]; x[5] = d.TB.xnm[fx];
        }
        if (hy) {
            y[0] = d.TB.ybt[fy]; y[1] = d.TB.ybx[fy]; y[2] = d.TB.ygam[fy];
            y[3] = d.TB.yib[fy]; y[4] = d.TB.ynp[fy]; y[5] = d.TB.ynm[fy];
        }
        if (hv) d.TA.vbt[v] = a;
        if (hc) {
            d.TA.btc[g] = c[0]; d.TA.cbx[g] = c[1]; d.TA.cby[g] = c[2]; d.TA.cgam[g] = c[3];
            d.TA.pxu[g] = c[4]; d.TA.pxv[g] = c[5]; d.TA.pyv[g] = c[6];
        }
        if (hx) {
            d.TA.xbt[fx] = x[0]; d.TA.xby[fx] = x[1]; d.TA.xgam[fx] = x[2];
            d.TA.xib[fx] = x[3]; d.TA.xnp[fx] = x[4]; d.TA.xnm[fx] = x[5];
        }
        if (hy) {
            d.TA.ybt[fy] = y[0]; d.TA.ybx[fy] = y[1]; d.TA.ygam[fy] = y[2];
            d.TA.yib[fy] = y[3]; d.TA.ynp[fy] = y[4]; d.TA.ynm[fy] = y[5];
        }
    }
}

// The accepted step's bed becomes the current one (CopySolutionData
// intermed3 -> tileContainer, src/TimeStepper.f90:263), in (tile, part) items
// over a grid sized for the active set the host last saw.
__global__ __launch_bounds__(256) void commit_kernel(const Dev* __restrict__ dp, const Params* __restrict__ pp) {
    const auto& d = const_arg(dp);
    const Params& p = *pp;
    const auto* c = ctrl_view(d);
    if (!c->commit) return;
    if (threadIdx.x == 0) count_work(d.ctrl, KW_COMMIT);
    const int n = c->n_active * TILE_PARTS;
    for (int it = blockIdx.x; it < n; it += gridDim.x)
        commit_part(d, p, d.active[it / TILE_PARTS], it % TILE_PARTS);
}

// ---------------------------------------------------------------------------
// batches with morphodynamic scenarios: the morphodynamic operator and the
// second hydraulic operator walk the work lists of the morphodynamic slots
// (Batch::off2: their 16 x 16 sub-blocks, Batch::off3: their (tile, part)
// items), the redistribution takes one workgroup per slot
// ---------------------------------------------------------------------------

// The slot of work item w of a list whose slot s starts at off[s].
__device__ __forceinline__ int batch_slot(const int* off, int nslot, int w) {
    int lo = 0, hi = nslot - 1;   // last slot with off <= w
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (off[mid] <= w) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int S>
__global__ __launch_bounds__(morpho_nt<B>()) void batch_morpho_kernel(Batch bt) {
    __shared__ int sh_slot;
    const int total = bt.off2[bt.nslot];
    const int m = gridDim.x / 8, x = blockIdx.x % 8;
    const int per = (total + 7) / 8;
    const int end = min((x + 1) * per, total);
    for (int w = x * per + (int)(blockIdx.x / 8); w < end; w += m) {
        if (threadIdx.x == 0) sh_slot = batch_slot(bt.off2, bt.nslot, w);
        __syncthreads();
        const int sl = __builtin_amdgcn_readfirstlane(sh_slot);
        morpho_pass<S, B, false>(const_arg(bt.dev + sl), generic(const_arg(bt.par + sl)), 0, w - bt.off2[sl]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(RD_THREADS) void batch_redistribute_kernel(Batch bt) {
    const int s = blockIdx.x;
    if (s >= bt.nslot || !bt.on[s] || !bt.dev[s].morpho) return;
    redistribute_body(const_arg(bt.dev + s), generic(const_arg(bt.par + s)));
}

// kind 0: the face topography of the new bed; 1: the accepted bed's commit
template <int KIND>
__global__ __launch_bounds__(256) void batch_tile_part_kernel(Batch bt) {
    __shared__ int sh_slot, sh_go;
    const int total = bt.off3[bt.nslot];
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        if (threadIdx.x == 0) {
            const int sl = batch_slot(bt.off3, bt.nslot, w);
            const auto& d = const_arg(bt.dev + sl);
            sh_slot = sl;
            sh_go = KIND == 0 ? (face_topo_due(d, generic(const_arg(bt.par + sl))) ? 1 : 0)
                              : (ctrl_view(d)->commit ? 1 : 0);
        }
        __syncthreads();
        const int sl = __builtin_amdgcn_readfirstlane(sh_slot);
        if (__builtin_amdgcn_readfirstlane(sh_go)) {
            const auto& d = const_arg(bt.dev + sl);
            const Params& p = generic(const_arg(bt.par + sl));
            const int it = w - bt.off3[sl];
            if (KIND == 0) face_topo_item(d, p, it);
            else commit_part(d, p, d.active[it / TILE_PARTS], it % TILE_PARTS);
        }
        __syncthreads();
    }
}

}  // namespace kst
