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
