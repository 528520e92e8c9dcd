// kst_physics.hpp -- pointwise physics of the time step, device side.
//
// Every function keeps the reference's evaluation order (Fortran evaluates
// a*b*c as (a*b)*c and -a*b/c as -((a*b)/c)); the library is built with
// -ffp-contract=off so the FP64 path reproduces the reference bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#include "../../include/kestrel_step.h"

#pragma clang fp contract(off)

namespace kst {

constexpr double kEps = DBL_EPSILON;   // Fortran epsilon(1.0_wp)
constexpr double kHuge = DBL_MAX;      // Fortran huge(1.0_wp)

// Parameters the kernels read, packed for a kernel argument.
struct Params {
    int nx, ny, ntx, nty, ntiles, NC, NFX, NFY;
    int limiter, geometric, drag, fswitch, tile_buffer, track_maxima, n_sources;
    int erosion, deposition, ero_transition, morpho_damping;
    double dx, dy, dxr, dyr, x_size, y_size;
    double g, rhow, rhos, chezy, coulomb, manning;
    double pmin, pmax, pint, pbeta, ebst, ekap, egam;
    double ero_rate, ero_rate_granular, ero_depth, ero_hcrit;
    double vrate, vvalue, max_pack, porosity, diam, nu;
    double cfl, max_dt, hthr, diff_ts;
    double gred, ws0, nsettling, crit_shields;
    // PouliquenStaticSlope (pmin + t1) / (1 - pmin t1), t1 = tan(1 degree),
    // of GranularErosion (src/Closures.f90:626-627), computed on the host
    double gran_stat;
    // boundary conditions (KST_BC_*); with Dirichlet ones the flow state of
    // the ghost tiles on the domain edge (SetDomainBoundaryData,
    // src/UpdateTiles.f90:595-643): depth, velocities, solids fraction and
    // their density
    int bc, periodic;
    double bc_hn, bc_u, bc_v, bc_psi, bc_rho;
    // a 1D grid (isOneD: nYtiles * nYpertile = 1, src/DomainSettings.f90:216-220)
    int oned;
};

__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return (a > b) ? a : b; }

// Limiters (src/Limiters.f90:83-184).
__device__ __forceinline__ double limiter(int kind, double a, double b) {
    switch (kind) {
    case KST_LIMITER_MINMOD1:
        if (a * b <= 0.0) return 0.0;
        return (a > 0.0) ? dmin(a, b) : dmax(a, b);
    case KST_LIMITER_MINMOD2: {
        const double th = 1.3;
        if (a * b <= 0.0) return 0.0;
        if (a > 0.0) return dmin(th * a, dmin(th * b, 0.5 * (a + b)));
        return dmax(th * a, dmax(th * b, 0.5 * (a + b)));
    }
    case KST_LIMITER_VANALBADA: {
        double den = a * a + b * b;
        if (den == 0.0) return 0.0;
        return (a * a * b + a * b * b) / den;
    }
    case KST_LIMITER_WENO: {
        const double eps = 1.0e-6;
        double ea = a * a + eps, eb = b * b + eps;
        double wa = 1.0 / (ea * ea), wb = 1.0 / (eb * eb);
        return (wa * a + wb * b) / (wa + wb);
    }
    default:
        return 0.5 * (a + b);
    }
}

// GeometricCorrectionFactor (src/Closures.f90:269-315).
__device__ __forceinline__ double gamma_of(int geometric, double bx, double by) {
    if (!geometric) return 1.0;
    return sqrt(1.0 + bx * bx + by * by);
}

// ComputeHn (src/Closures.f90:158-171).
__device__ __forceinline__ double compute_hn(double w, double b0, double bt, double gam) {
    double hn = -bt;
    hn = hn + (w - b0);
    return hn * gam;
}

// KahanAdd (src/utilities.f90:439-448).
struct Kahan {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x) {
        double y = x - c;
        double t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// FlowSquaredSpeedSlopeAligned (src/Closures.f90:178-205), 2D.
__device__ __forceinline__ double speed2(int geometric, double u, double v, double bx, double by) {
    double m;
    if (geometric) {
        m = u * u * (1.0 + bx * bx);
        m = m + v * v * (1.0 + by * by) + 2.0 * bx * by * u * v;
    } else {
        m = u * u;
        m = m + v * v;
    }
    return m;
}

__device__ __forceinline__ double density(const Params& p, double psi) {
    return p.rhow + (p.rhos - p.rhow) * psi;
}

// Tile di, dj (each -1..1) tiles from t.  Without periodic boundaries plain
// arithmetic: the tiles an active tile reads lie inside the grid, since edge
// tiles are never activated then (AddTile, src/UpdateTiles.f90:56-78).  With
// them the neighbours wrap around the grid (SetPeriodicBoundaryConds,
// src/UpdateTiles.f90:521-566; a diagonal tile is the wrapped neighbour's
// wrapped neighbour, as the reference reaches it through North/South of the
// East/West tile).
__device__ __forceinline__ int nb_tile(const Params& p, int t, int di, int dj) {
    if (!p.periodic) return t + di + dj * p.ntx;
    int ti = t % p.ntx + di, tj = t / p.ntx + dj;
    ti += ti < 0 ? p.ntx : (ti >= p.ntx ? -p.ntx : 0);
    tj += tj < 0 ? p.nty : (tj >= p.nty ? -p.nty : 0);
    return ti + tj * p.ntx;
}

// A tile in the first or last column or row of the grid (the domain-edge
// test of AddTile and UpdateGhostTileBCs, src/UpdateTiles.f90:56-78,571-585).
__device__ __forceinline__ bool edge_tile(const Params& p, int t) {
    const int ti = t % p.ntx, tj = t / p.ntx;
    return ti == 0 || ti == p.ntx - 1 || (p.nty > 1 && (tj == 0 || tj == p.nty - 1));
}

// A ghost tile holding the Dirichlet boundary data, whose derived fields are
// the boundary values (no other ghost has any: SetDefaultTileData zeroes them).
__device__ __forceinline__ bool dirichlet_ghost(const Params& p, int t) {
    return p.bc == KST_BC_DIRICHLET && edge_tile(p, t);
}

// fswitch family (src/Closures.f90:797-915).
__device__ inline double fswitch(const Params& p, double psi) {
    switch (p.fswitch) {
    case KST_SWITCH_TANH:
        return 0.5 * (1.0 + tanh(p.vrate * (psi - p.vvalue)));
    case KST_SWITCH_RAT3: {
        double a = p.vvalue - 1.5 / p.vrate;
        double b = p.vvalue + 1.5 / p.vrate;
        if (psi <= a) return 0.0;
        if (psi >= b) return 1.0;
        double x = (psi - a) / (b - a);
        return (x * x * x) / ((1 - x) * (1 - x) * (1 - x) + x * x * x);
    }
    case KST_SWITCH_COS: {
        const double pi = 3.141592653589793;
        double a = p.vvalue - 0.25 * pi / p.vrate;
        double b = p.vvalue + 0.25 * pi / p.vrate;
        if (psi <= a) return 0.0;
        if (psi >= b) return 1.0;
        double x = (psi - a) / (b - a);
        return 0.5 * (1.0 - cos(pi * x));
    }
    case KST_SWITCH_EQUAL: return 0.5;
    case KST_SWITCH_ZERO: return 0.0;
    case KST_SWITCH_ONE: return 1.0;
    case KST_SWITCH_LINEAR: return psi / p.max_pack;
    default: return (psi < p.vvalue) ? 0.0 : 1.0;
    }
}

// PouliquenFrictionCoefficient (src/Closures.f90:436-458).
__device__ inline double pouliquen_mu(const Params& p, double gcos, double hn, double modu) {
    if (hn > p.hthr) {
        double fr = modu / sqrt(gcos * hn);
        double I = fr * p.diam / hn;
        return p.pmin + (p.pmax - p.pmin) * I / (p.pbeta + I);
    }
    return p.pmin;
}

__device__ inline double drag_chezy(const Params& p, double u, double v, double bx, double by) {
    return p.chezy * speed2(p.geometric, u, v, bx, by);
}
__device__ inline double drag_coulomb(const Params& p, double hn, double gam) {
    double g = p.g / gam;
    return p.coulomb * g * hn;
}
__device__ inline double drag_pouliquen(const Params& p, double hn, double u, double v, double bx,
                                        double by, double gam) {
    double g = p.g / gam;
    double m2 = speed2(p.geometric, u, v, bx, by);
    double mu = pouliquen_mu(p, g, hn, sqrt(m2));
    if (m2 > 0) return mu * g * hn;
    return 0.0;
}

// DragClosure family (src/Closures.f90:365-558).
__device__ inline double drag(const Params& p, double hn, double u, double v, double psi, double bx,
                              double by, double gam) {
    switch (p.drag) {
    case KST_DRAG_CHEZY: return drag_chezy(p, u, v, bx, by);
    case KST_DRAG_COULOMB: return drag_coulomb(p, hn, gam);
    case KST_DRAG_VOELLMY: return drag_chezy(p, u, v, bx, by) + drag_coulomb(p, hn, gam);
    case KST_DRAG_POULIQUEN: return drag_pouliquen(p, hn, u, v, bx, by, gam);
    case KST_DRAG_EDWARDS2019: {
        double gperp = p.g / gam;
        double modu = sqrt(speed2(p.geometric, u, v, bx, by));
        double fr = modu / sqrt(gperp * hn);
        double mu1 = p.pmin, mu2 = p.pmax, mu3 = p.pint, L = p.diam, f;
        if (fr > p.ebst) {
            f = mu1 + (mu2 - mu1) / (1.0 + hn * p.pbeta / (L * (fr + p.egam)));
        } else {
            f = (pow(fr / p.ebst, p.ekap)) *
                    (mu1 + (mu2 - mu1) / (1.0 + hn * p.pbeta / (L * (p.ebst + p.egam))) - mu3 -
                     (mu2 - mu1) / (1.0 + hn / L)) +
                mu3 + (mu2 - mu1) / (1.0 + hn / L);
        }
        return f * gperp * hn;
    }
    case KST_DRAG_MANNING: {
        double g = p.g / gam;
        if (hn > p.hthr) return g * p.manning * p.manning / pow(hn, 1.0 / 3.0);
        return 0.0;
    }
    default: {
        double ch = drag_chezy(p, u, v, bx, by);
        double po = drag_pouliquen(p, hn, u, v, bx, by, gam);
        double fc = fswitch(p, psi);
        return ch * (1.0 - fc) + po * fc;
    }
    }
}

// ---- morphodynamic closures ------------------------------------------------

// Cell state the erosion/deposition closures read (fields Hn, u, v, psi, bt,
// dbdx, dbdy of the reference's u vector).
struct CellState {
    double hn, u, v, psi, bt, bx, by;
};

// FlowShieldsNumber / FlowParticleSpeed (src/Closures.f90:208-241).
__device__ inline double shields_number(const Params& p, const CellState& c) {
    const double gred = p.gred / gamma_of(p.geometric, c.bx, c.by);
    const double m2 = speed2(p.geometric, c.u, c.v, c.bx, c.by);
    const double chezy = p.chezy * m2;
    return chezy / (gred * p.diam);
}
__device__ inline double particle_speed(const Params& p, const CellState& c) {
    const double gred = p.gred / gamma_of(p.geometric, c.bx, c.by);
    return sqrt(gred * p.diam);
}

// FluidErosion, GranularErosion (src/Closures.f90:588-655).
__device__ inline double erosion_fluid(const Params& p, const CellState& c) {
    const double sh = shields_number(p, c);
    if (sh > p.crit_shields) {
        double e = p.ero_rate * (sh - p.crit_shields);
        return e * particle_speed(p, c);
    }
    return 0.0;
}
__device__ inline double erosion_granular(const Params& p, const CellState& c) {
    const double m2 = speed2(p.geometric, c.u, c.v, c.bx, c.by);
    const double mu1 = p.pmin;
    const double stat = p.gran_stat;
    const double gcos = p.g / gamma_of(p.geometric, c.bx, c.by);
    const double mu = pouliquen_mu(p, gcos, c.hn, sqrt(m2));
    // (Hn/25/d)**2.0_wp: the square, as gfortran evaluates a real power 2
    const double r = c.hn / 25.0 / p.diam;
    const double mun = mu1 + (stat - mu1) / (1.0 + r * r);
    if (mu > mun) {
        double e = p.ero_rate_granular * (mu - mun);
        return e * particle_speed(p, c);
    }
    return 0.0;
}

// ErosionClosure family (src/Closures.f90:566-675).
__device__ inline double erosion_closure(const Params& p, const CellState& c) {
    switch (p.erosion) {
    case KST_EROSION_SIMPLE: {
        double e = p.ero_rate * shields_number(p, c);
        return e * particle_speed(p, c);
    }
    case KST_EROSION_FLUID: return erosion_fluid(p, c);
    case KST_EROSION_GRANULAR: return erosion_granular(p, c);
    case KST_EROSION_MIXED: {
        const double fc = fswitch(p, c.psi);
        const double fl = erosion_fluid(p, c);
        const double gr = erosion_granular(p, c);
        return (1.0 - fc) * fl + fc * gr;
    }
    default: return 0.0;
    }
}

// ErosionTransition family (src/Closures.f90:684-732).
__device__ inline double erosion_transition(const Params& p, double bt) {
    switch (p.ero_transition) {
    case KST_TRANSITION_SMOOTH: return 0.5 * (1.0 + tanh(1e5 * (bt + p.ero_depth)));
    case KST_TRANSITION_STEP: return (bt < -p.ero_depth) ? 0.0 : 1.0;
    default: return 1.0;
    }
}

// DepositionClosure family (src/Closures.f90:320-356).
__device__ inline double deposition_closure(const Params& p, double psi) {
    switch (p.deposition) {
    case KST_DEPOSITION_SIMPLE: return psi * (1.0 - psi / p.max_pack);
    case KST_DEPOSITION_SPEARMAN_MANNING: {
        const double a = 2.7 - 0.15 * p.nsettling;
        const double b = 0.62 * p.nsettling - 1.46;
        return psi * pow(1.0 - psi, a) * pow(1.0 - psi / p.max_pack, b);
    }
    default: return 0.0;
    }
}

// MorphoDamping family (src/Closures.f90:744-791).
__device__ inline double morpho_damping(const Params& p, double hn) {
    switch (p.morpho_damping) {
    case KST_DAMPING_TANH: return 0.5 * (1.0 + tanh(10.0 * log(hn / p.ero_hcrit)));
    case KST_DAMPING_RAT3: {
        const double hc = p.ero_hcrit;
        if (hn < hc) return 0.0;
        if (hn > 2.0 * hc) return 1.0;
        const double t = hn / hc - 1.0;
        return (t * t * t) / ((1.0 - t) * (1.0 - t) * (1.0 - t) + t * t * t);
    }
    default: return 1.0;
    }
}

// ErosionDepositionTerms, Erosion, Deposition (src/Equations.f90:385-448):
// returns E - D.
// E - D for the reference's default laws (Mixed erosion, Spearman-Manning
// deposition, smooth erosion transition, tanh damping, tanh switch) in
// straight-line code: the same operations as the general form below, with
// its data-dependent branches as selects, so that the independent
// transcendental chains (three tanh, a log, two pow, square roots) overlap
// instead of running one after another.
__device__ __forceinline__ double emd_default_laws(const Params& p, const CellState& c) {
    const double gam = gamma_of(p.geometric, c.bx, c.by);
    const double m2 = speed2(p.geometric, c.u, c.v, c.bx, c.by);
    const double gred = p.gred / gam;
    const double ps = sqrt(gred * p.diam);                          // FlowParticleSpeed
    // fluid erosion (FlowShieldsNumber)
    const double sh = (p.chezy * m2) / (gred * p.diam);
    const double efl = (sh > p.crit_shields) ? (p.ero_rate * (sh - p.crit_shields)) * ps : 0.0;
    // granular erosion (PouliquenFrictionCoefficient at g cos = g / gamma)
    const double gcos = p.g / gam;
    const double fr = sqrt(m2) / sqrt(gcos * c.hn);
    const double I = fr * p.diam / c.hn;
    const double mu = (c.hn > p.hthr) ? p.pmin + (p.pmax - p.pmin) * I / (p.pbeta + I) : p.pmin;
    const double r = c.hn / 25.0 / p.diam;
    const double mun = p.pmin + (p.gran_stat - p.pmin) / (1.0 + r * r);
    const double egr = (mu > mun) ? (p.ero_rate_granular * (mu - mun)) * ps : 0.0;
    const double fc = 0.5 * (1.0 + tanh(p.vrate * (c.psi - p.vvalue)));
    double e = (1.0 - fc) * efl + fc * egr;
    e = e * (0.5 * (1.0 + tanh(1e5 * (c.bt + p.ero_depth))));
    const double a = 2.7 - 0.15 * p.nsettling;
    const double b = 0.62 * p.nsettling - 1.46;
    const double sm = c.psi * pow(1.0 - c.psi, a) * pow(1.0 - c.psi / p.max_pack, b);
    const double alpha = (c.psi >= p.max_pack) ? 0.0 : ((c.psi > 0.0) ? sm : 0.0);
    const double dep = p.ws0 * alpha;
    const double damp = 0.5 * (1.0 + tanh(10.0 * log(c.hn / p.ero_hcrit)));
    return e * damp - dep * damp;
}

__device__ inline double erosion_minus_deposition(const Params& p, const CellState& c) {
    if (p.erosion == KST_EROSION_MIXED && p.deposition == KST_DEPOSITION_SPEARMAN_MANNING &&
        p.ero_transition == KST_TRANSITION_SMOOTH && p.morpho_damping == KST_DAMPING_TANH &&
        p.fswitch == KST_SWITCH_TANH)
        return emd_default_laws(p, c);
    double e = erosion_closure(p, c);
    e = e * erosion_transition(p, c.bt);
    double alpha;
    if (c.psi >= p.max_pack) alpha = 0.0;
    else if (c.psi > 0.0) alpha = deposition_closure(p, c.psi);
    else alpha = 0.0;
    const double dep = p.ws0 * alpha;
    const double damp = morpho_damping(p, c.hn);
    return e * damp - dep * damp;
}

// Square root of the gravity-wave speed term of the wave speeds
// (src/Equations.f90:249-381).
__device__ __forceinline__ double wave_root(const Params& p, double hn, double bperp, double gam) {
    if (hn <= 0.0) hn = 0.0;
    if (p.geometric) return sqrt(p.g * hn * (1.0 + bperp * bperp) / (gam * gam * gam));
    return sqrt(p.g * hn);
}

// Positive doubles order like their bit patterns, so a CFL minimum can be
// reduced with an integer atomicMin.
__device__ __forceinline__ unsigned long long dbits(double x) {
    return (unsigned long long)__double_as_longlong(x);
}
__device__ __forceinline__ double bitsd(unsigned long long b) {
    return __longlong_as_double((long long)b);
}

}  // namespace kst
