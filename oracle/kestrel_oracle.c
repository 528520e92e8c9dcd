/*
 * kestrel_oracle.c -- CPU restatement of Kestrel's finite-volume step
 * (hydraulic operator, SSP-RK3 stepper with semi-implicit drag, CFL time
 * step, wet-tile activation, running maxima).
 *
 * TEST INFRASTRUCTURE ONLY: the oracle.  It is loaded only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
 *
 * Every routine restates the reference routine cited beside it, operation by
 * operation, in the reference's evaluation order (Fortran evaluates a*b*c as
 * (a*b)*c and -a*b/c as -((a*b)/c)); build it with -ffp-contract=off so no
 * multiply-add is fused.  The data model is the reference's: tiles of
 * nx*ny cells holding 13 fields, face reconstructions, fluxes and RHS
 * arrays, a tileContainer plus four intermed containers refreshed by deep
 * copies, a sorted active-tile list and ghost tiles holding frozen data.
 *
 * Pinning: the hydraulic and morphodynamic RHS, topography and
 * redistribution equal the reference's own compiled Fortran (oracle/_ref)
 * bit for bit on the committed vectors; the stepper, activation, maxima and
 * ko_volume are pinned by conservation (the reference's 1e-10 check), exact
 * sums, closed forms, symmetry, tile-layout independence and planted
 * mistakes.  Parity unpinned: the last-ulp behaviour of tanh, log, pow and
 * cos (glibc here, the ROCm device library on the GPU) in the closures that
 * call them; decks using them are compared to 1e-9 relative.
 *
 * Field indices (src/main.f90:76-103), 0-based here.
 */
#include "kestrel_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef KO_PERTURB_LIBM
/* Conditioning probe (liboracle_pert.so): the transcendental functions the
 * closures evaluate per cell return their result moved one ulp up.  A run
 * compared with the plain oracle's measures how far last-bit differences
 * between two libm implementations (the host's and the device's) carry each
 * field; where that spread exceeds the parity bar, the device is held to it.
 * The constants pi = 4 atan(1) and tan(pi/180) are left exact. */
static double ko_up(double x) { return nextafter(x, INFINITY); }
static double ko_tanh(double x) { return ko_up(tanh(x)); }
static double ko_cos(double x) { return ko_up(cos(x)); }
static double ko_pow(double x, double y) { return ko_up(pow(x, y)); }
static double ko_log(double x) { return ko_up(log(x)); }
#define tanh ko_tanh
#define cos ko_cos
#define pow ko_pow
#define log ko_log
#endif

enum { W = 0, RHU = 1, RHV = 2, HNPSI = 3, HN = 4, UU = 5, VV = 6, PSI = 7, RHO = 8,
       B0 = 9, BT = 10, DBDX = 11, DBDY = 12, NF = 13 };

static const int IFLUX[4] = {W, RHU, RHV, HNPSI};
static const int IDESING[5] = {HN, UU, VV, PSI, RHO};

#define EPS_WP DBL_EPSILON /* Fortran epsilon(1.0_wp) */
#define HUGE_WP DBL_MAX    /* Fortran huge(1.0_wp) */

/* Per-container tile data (src/Grid.f90:57-121). */
typedef struct {
    int alloc;
    double *u, *ulx, *uly;            /* [13][ny][nx] */
    double *upx, *umx;                /* [13][ny][nx+1] */
    double *upy, *umy;                /* [13][ny+1][nx] */
    double *hx, *gx, *px;             /* [4][ny][nx+1] */
    double *hy, *gy, *py;             /* [4][ny+1][nx] */
    double *dte, *dti;                /* [4][ny][nx] */
    double *dtbt;                     /* [ny+1][nx+1] */
    double *b0, *bt;                  /* [ny+1][nx+1] */
    double *emd;                      /* [ny][nx] */
} tdata;

/* Container-independent tile metadata. */
typedef struct {
    int on, ghost, src, touched;
    int nb[4];   /* W, E, N, S  (src/UpdateTiles.f90:494-497) */
    int cr[4];   /* SW, SE, NW, NE (src/UpdateTiles.f90:498-501) */
    double *x, *y, *xv, *yv;
    double *hnmax, *umax, *emax, *dmax, *psimax; /* [2][ny][nx] */
    double *tfirst;
    double xmin, xmax, ymin, ymax;
} tmeta;

enum { C_TC = 0, C_I0 = 1, C_I1 = 2, C_I2 = 3, C_I3 = 4, NCONT = 5 };

struct ko_sim {
    kst_config cfg;
    kst_flux_source* src;     /* private copy of the flux sources */
    double** src_arrays;
    int nx, ny, ntx, nty, ntiles, NX, NY;
    int NC, NFX, NFY, NV;
    int oned;                 /* isOneD: nYtiles * nYpertile = 1 (src/DomainSettings.f90:216-220) */
    double dxr, dyr;
    tmeta* meta;
    tdata* cont[NCONT];
    int* active; int nactive;
    int* ghosts; int nghost;
    double t, t0, dt;
    double* gb0;              /* global vertex b0 */
    char err[256];
    int status;
    /* probe of the first redistribution of a run (for pinning RedistributeGrid) */
    int probe_on, probe_done;
    tdata* probe[3];          /* I0, I3 before and I3 after redistribution */
    void* probe_list; int probe_n;
    uint8_t* probe_ts;        /* tile states at the probe */
    int64_t redist_cells;     /* RedistributeCell calls that succeeded */
    int64_t redist_ghost;     /* their writes into ghost tiles (vertex copies, 3x3 cells) */
    double* steplog;          /* (t, dt) of each accepted step while logging (ko_step_log) */
    int64_t steplog_n, steplog_cap;
    double probe_dt, probe_dt_pending;
};

#define NCw(s) ((s)->NC)

static int fail(ko_sim* s, int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(s->err, sizeof s->err, fmt, ap);
    va_end(ap);
    s->status = code;
    return code;
}

/* ------------------------------------------------------------------ */
/* Pointwise physics                                                  */
/* ------------------------------------------------------------------ */

static double fmin_(double a, double b) { return (a < b) ? a : b; }
static double fmax_(double a, double b) { return (a > b) ? a : b; }

/* Limiters (src/Limiters.f90:83-184). */
static double limiter(const kst_config* c, double a, double b) {
    switch (c->limiter) {
    case KST_LIMITER_MINMOD1:
        if (a * b <= 0.0) return 0.0;
        return (a > 0.0) ? fmin_(a, b) : fmax_(a, b);
    case KST_LIMITER_MINMOD2: {
        const double th = 1.3;
        if (a * b <= 0.0) return 0.0;
        if (a > 0.0) return fmin_(th * a, fmin_(th * b, 0.5 * (a + b)));
        return fmax_(th * a, fmax_(th * b, 0.5 * (a + b)));
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

/* GeometricCorrectionFactor (src/Closures.f90:269-315). */
static double gamma_of(const kst_config* c, double bx, double by) {
    if (!c->geometric_factors) return 1.0;
    return sqrt(1.0 + bx * bx + by * by);
}
static double gam_v(const kst_config* c, const double* uv, int stride) {
    return gamma_of(c, uv[DBDX * stride], uv[DBDY * stride]);
}

/* ComputeHn (src/Closures.f90:158-171). */
static double compute_hn(double w, double b0, double bt, double gam) {
    double hn = -bt;
    hn = hn + (w - b0);
    hn = hn * gam;
    return hn;
}

/* KahanAdd / KahanSum (src/utilities.f90:418-448). */
static double kahan_sum(const double* x, int n) {
    double s = 0.0, c = 0.0;
    for (int i = 0; i < n; ++i) {
        double y = x[i] - c;
        double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    return s;
}

/* FlowSquaredSpeedSlopeAligned (src/Closures.f90:178-205); on a 1D grid the
 * x part only. */
static double speed2(const kst_config* c, double u, double v, double bx, double by) {
    double m;
    if (c->geometric_factors) m = u * u * (1.0 + bx * bx);
    else m = u * u;
    if (c->ny_tiles * c->ny_per_tile == 1) return m;
    if (c->geometric_factors) m = m + v * v * (1.0 + by * by) + 2.0 * bx * by * u * v;
    else m = m + v * v;
    return m;
}

static double density(const kst_config* c, double psi) { return c->rhow + (c->rhos - c->rhow) * psi; }

/* fswitch family (src/Closures.f90:797-915). */
static double fswitch(const kst_config* c, double psi) {
    switch (c->fswitch) {
    case KST_SWITCH_TANH:
        return 0.5 * (1.0 + tanh(c->voellmy_switch_rate * (psi - c->voellmy_switch_value)));
    case KST_SWITCH_RAT3: {
        double a = c->voellmy_switch_value - 1.5 / c->voellmy_switch_rate;
        double b = c->voellmy_switch_value + 1.5 / c->voellmy_switch_rate;
        if (psi <= a) return 0.0;
        if (psi >= b) return 1.0;
        double x = (psi - a) / (b - a);
        return (x * x * x) / ((1 - x) * (1 - x) * (1 - x) + x * x * x);
    }
    case KST_SWITCH_COS: {
        const double pi = 4.0 * atan(1.0);
        double a = c->voellmy_switch_value - 0.25 * pi / c->voellmy_switch_rate;
        double b = c->voellmy_switch_value + 0.25 * pi / c->voellmy_switch_rate;
        if (psi <= a) return 0.0;
        if (psi >= b) return 1.0;
        double x = (psi - a) / (b - a);
        return 0.5 * (1.0 - cos(pi * x));
    }
    case KST_SWITCH_EQUAL: return 0.5;
    case KST_SWITCH_ZERO: return 0.0;
    case KST_SWITCH_ONE: return 1.0;
    case KST_SWITCH_LINEAR: return psi / c->max_pack;
    default: return (psi < c->voellmy_switch_value) ? 0.0 : 1.0;
    }
}

/* PouliquenFrictionCoefficient (src/Closures.f90:436-458). */
static double pouliquen_mu(const kst_config* c, double gcos, double hn, double modu) {
    double mu1 = c->pouliquen_min, mu2 = c->pouliquen_max, beta = c->pouliquen_beta;
    if (hn > c->height_threshold) {
        double fr = modu / sqrt(gcos * hn);
        double I = fr * c->solid_diameter / hn;
        return mu1 + (mu2 - mu1) * I / (beta + I);
    }
    return mu1;
}

/* Cell state view for the closures. */
typedef struct { double hn, u, v, psi, bt, bx, by; } cstate;

static double drag_chezy(const kst_config* c, const cstate* s) {
    return c->chezy_co * speed2(c, s->u, s->v, s->bx, s->by);
}
static double drag_coulomb(const kst_config* c, const cstate* s) {
    double gam = gamma_of(c, s->bx, s->by);
    double g = c->g / gam;
    return c->coulomb_co * g * s->hn;
}
static double drag_pouliquen(const kst_config* c, const cstate* s) {
    double gam = gamma_of(c, s->bx, s->by);
    double g = c->g / gam;
    double m2 = speed2(c, s->u, s->v, s->bx, s->by);
    double mu = pouliquen_mu(c, g, s->hn, sqrt(m2));
    if (m2 > 0) return mu * g * s->hn;
    return 0.0;
}

/* DragClosure family (src/Closures.f90:365-558). */
static double drag(const kst_config* c, const cstate* s) {
    switch (c->drag) {
    case KST_DRAG_CHEZY: return drag_chezy(c, s);
    case KST_DRAG_COULOMB: return drag_coulomb(c, s);
    case KST_DRAG_VOELLMY: return drag_chezy(c, s) + drag_coulomb(c, s);
    case KST_DRAG_POULIQUEN: return drag_pouliquen(c, s);
    case KST_DRAG_EDWARDS2019: {
        double hn = s->hn;
        double gam = gamma_of(c, s->bx, s->by);
        double gperp = c->g / gam;
        double modu = sqrt(speed2(c, s->u, s->v, s->bx, s->by));
        double fr = modu / sqrt(gperp * hn);
        double mu1 = c->pouliquen_min, mu2 = c->pouliquen_max, mu3 = c->pouliquen_intermediate;
        double beta = c->pouliquen_beta, bst = c->edwards_betastar, kap = c->edwards_kappa;
        double cg = c->edwards_gamma, L = c->solid_diameter, f;
        if (fr > bst) {
            f = mu1 + (mu2 - mu1) / (1.0 + hn * beta / (L * (fr + cg)));
        } else {
            f = (pow(fr / bst, kap)) *
                    (mu1 + (mu2 - mu1) / (1.0 + hn * beta / (L * (bst + cg))) - mu3 -
                     (mu2 - mu1) / (1.0 + hn / L)) +
                mu3 + (mu2 - mu1) / (1.0 + hn / L);
        }
        return f * gperp * hn;
    }
    case KST_DRAG_MANNING: {
        double gam = gamma_of(c, s->bx, s->by);
        double g = c->g / gam;
        if (s->hn > c->height_threshold)
            return g * c->manning_co * c->manning_co / pow(s->hn, 1.0 / 3.0);
        return 0.0;
    }
    default: { /* Variable */
        double ch = drag_chezy(c, s);
        double po = drag_pouliquen(c, s);
        double fc = fswitch(c, s->psi);
        return ch * (1.0 - fc) + po * fc;
    }
    }
}

/* Shields number and particle speed (src/Closures.f90:208-241). */
static double shields_number(const kst_config* c, const cstate* s) {
    double gred = c->gred / gamma_of(c, s->bx, s->by);
    double m2 = speed2(c, s->u, s->v, s->bx, s->by);
    double chezy = c->chezy_co * m2;
    return chezy / (gred * c->solid_diameter);
}
static double particle_speed(const kst_config* c, const cstate* s) {
    double gred = c->gred / gamma_of(c, s->bx, s->by);
    return sqrt(gred * c->solid_diameter);
}

static double erosion_fluid(const kst_config* c, const cstate* s) {
    double sh = shields_number(c, s);
    if (sh > c->critical_shields) {
        double e = c->ero_rate * (sh - c->critical_shields);
        return e * particle_speed(c, s);
    }
    return 0.0;
}
static double erosion_granular(const kst_config* c, const cstate* s) {
    const double pi = 4.0 * atan(1.0);
    double hn = s->hn;
    double m2 = speed2(c, s->u, s->v, s->bx, s->by);
    double mu1 = c->pouliquen_min;
    double t1 = tan(pi / 180.0);
    double stat = (mu1 + t1) / (1.0 - mu1 * t1);
    double gcos = c->g / gamma_of(c, s->bx, s->by);
    double mu = pouliquen_mu(c, gcos, hn, sqrt(m2));
    /* (Hn/25/d)**2.0_wp (src/Closures.f90:631-632): gfortran evaluates a
       real power 2 as the square */
    double r = hn / 25.0 / c->solid_diameter;
    double mun = mu1 + (stat - mu1) / (1.0 + r * r);
    if (mu > mun) {
        double e = c->ero_rate_granular * (mu - mun);
        return e * particle_speed(c, s);
    }
    return 0.0;
}

/* ErosionClosure family (src/Closures.f90:566-675). */
static double erosion_closure(const kst_config* c, const cstate* s) {
    switch (c->erosion) {
    case KST_EROSION_SIMPLE: {
        double e = c->ero_rate * shields_number(c, s);
        return e * particle_speed(c, s);
    }
    case KST_EROSION_FLUID: return erosion_fluid(c, s);
    case KST_EROSION_GRANULAR: return erosion_granular(c, s);
    case KST_EROSION_MIXED: {
        double fc = fswitch(c, s->psi);
        double fl = erosion_fluid(c, s);
        double gr = erosion_granular(c, s);
        return (1.0 - fc) * fl + fc * gr;
    }
    default: return 0.0;
    }
}

/* ErosionTransition family (src/Closures.f90:684-732). */
static double erosion_transition(const kst_config* c, const cstate* s) {
    switch (c->ero_transition) {
    case KST_TRANSITION_SMOOTH: return 0.5 * (1.0 + tanh(1e5 * (s->bt + c->ero_depth)));
    case KST_TRANSITION_STEP: return (s->bt < -c->ero_depth) ? 0.0 : 1.0;
    default: return 1.0;
    }
}

/* DepositionClosure family (src/Closures.f90:320-356). */
static double deposition_closure(const kst_config* c, double psi) {
    switch (c->deposition) {
    case KST_DEPOSITION_SIMPLE: return psi * (1.0 - psi / c->max_pack);
    case KST_DEPOSITION_SPEARMAN_MANNING: {
        double a = 2.7 - 0.15 * c->nsettling;
        double b = 0.62 * c->nsettling - 1.46;
        return psi * pow(1.0 - psi, a) * pow(1.0 - psi / c->max_pack, b);
    }
    default: return 0.0;
    }
}

/* MorphoDamping family (src/Closures.f90:744-791). */
static double morpho_damping(const kst_config* c, double hn) {
    switch (c->morpho_damping) {
    case KST_DAMPING_TANH: return 0.5 * (1.0 + tanh(10.0 * log(hn / c->ero_critical_height)));
    case KST_DAMPING_RAT3: {
        double hc = c->ero_critical_height;
        if (hn < hc) return 0.0;
        if (hn > 2.0 * hc) return 1.0;
        double t = hn / hc - 1.0;
        return (t * t * t) / ((1.0 - t) * (1.0 - t) * (1.0 - t) + t * t * t);
    }
    default: return 1.0;
    }
}

/* ErosionDepositionTerms, Erosion, Deposition (src/Equations.f90:385-448). */
static void erosion_deposition(const kst_config* c, const cstate* s, double* E, double* D) {
    double e = erosion_closure(c, s);
    e = e * erosion_transition(c, s);
    double alpha;
    if (s->psi >= c->max_pack) alpha = 0.0;
    else if (s->psi > 0.0) alpha = deposition_closure(c, s->psi);
    else alpha = 0.0;
    double dep = c->ws0 * alpha;
    double damp = morpho_damping(c, s->hn);
    *E = e * damp;
    *D = dep * damp;
}

/* ------------------------------------------------------------------ */
/* Tile bookkeeping                                                   */
/* ------------------------------------------------------------------ */

static int in_bounds(const ko_sim* s, int k) { return k >= 0 && k < s->ntiles; }

static void tile_coords(const ko_sim* s, int k, int* gi, int* gj) {
    *gi = k % s->ntx;
    *gj = k / s->ntx;
}

/* OnDomainEdge (src/Grid.f90:310-323). */
static int on_domain_edge(const ko_sim* s, int k) {
    int gi, gj;
    tile_coords(s, k, &gi, &gj);
    int on = (gi == 0 || gi == s->ntx - 1);
    on = on || (s->nty > 1 && (gj == 0 || gj == s->nty - 1));
    return on;
}

static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

static void tdata_alloc(ko_sim* s, tdata* d) {
    if (d->alloc) return;
    int NC = s->NC, NFX = s->NFX, NFY = s->NFY, NV = s->NV;
    d->u = xcalloc(NF * NC, 8); d->ulx = xcalloc(NF * NC, 8); d->uly = xcalloc(NF * NC, 8);
    d->upx = xcalloc(NF * NFX, 8); d->umx = xcalloc(NF * NFX, 8);
    d->upy = xcalloc(NF * NFY, 8); d->umy = xcalloc(NF * NFY, 8);
    d->hx = xcalloc(4 * NFX, 8); d->gx = xcalloc(4 * NFX, 8); d->px = xcalloc(4 * NFX, 8);
    d->hy = xcalloc(4 * NFY, 8); d->gy = xcalloc(4 * NFY, 8); d->py = xcalloc(4 * NFY, 8);
    d->dte = xcalloc(4 * NC, 8); d->dti = xcalloc(4 * NC, 8);
    d->dtbt = xcalloc(NV, 8); d->b0 = xcalloc(NV, 8); d->bt = xcalloc(NV, 8);
    d->emd = xcalloc(NC, 8);
    d->alloc = 1;
}

static void tdata_free(tdata* d) {
    if (!d->alloc) return;
    free(d->u); free(d->ulx); free(d->uly); free(d->upx); free(d->umx); free(d->upy);
    free(d->umy); free(d->hx); free(d->gx); free(d->px); free(d->hy); free(d->gy);
    free(d->py); free(d->dte); free(d->dti); free(d->dtbt); free(d->b0); free(d->bt);
    free(d->emd);
    memset(d, 0, sizeof *d);
}

/* Deep copy of a TileType (intrinsic assignment, src/TimeStepper.f90:795-806). */
static void tdata_copy(ko_sim* s, tdata* dst, const tdata* src) {
    tdata_alloc(s, dst);
    int NC = s->NC, NFX = s->NFX, NFY = s->NFY, NV = s->NV;
    memcpy(dst->u, src->u, 8 * NF * NC); memcpy(dst->ulx, src->ulx, 8 * NF * NC);
    memcpy(dst->uly, src->uly, 8 * NF * NC);
    memcpy(dst->upx, src->upx, 8 * NF * NFX); memcpy(dst->umx, src->umx, 8 * NF * NFX);
    memcpy(dst->upy, src->upy, 8 * NF * NFY); memcpy(dst->umy, src->umy, 8 * NF * NFY);
    memcpy(dst->hx, src->hx, 8 * 4 * NFX); memcpy(dst->gx, src->gx, 8 * 4 * NFX);
    memcpy(dst->px, src->px, 8 * 4 * NFX);
    memcpy(dst->hy, src->hy, 8 * 4 * NFY); memcpy(dst->gy, src->gy, 8 * 4 * NFY);
    memcpy(dst->py, src->py, 8 * 4 * NFY);
    memcpy(dst->dte, src->dte, 8 * 4 * NC); memcpy(dst->dti, src->dti, 8 * 4 * NC);
    memcpy(dst->dtbt, src->dtbt, 8 * NV); memcpy(dst->b0, src->b0, 8 * NV);
    memcpy(dst->bt, src->bt, 8 * NV); memcpy(dst->emd, src->emd, 8 * NC);
}

/* Index helpers. */
#define CI(s, i, j) ((j) * (s)->nx + (i))
#define FXI(s, i, j) ((j) * ((s)->nx + 1) + (i))
#define FYI(s, i, j) ((j) * (s)->nx + (i))
#define VI(s, i, j) ((j) * ((s)->nx + 1) + (i))
#define UF(s, T, d, i, j) (T)->u[(d) * (s)->NC + CI(s, i, j)]
#define LX(s, T, d, i, j) (T)->ulx[(d) * (s)->NC + CI(s, i, j)]
#define LY(s, T, d, i, j) (T)->uly[(d) * (s)->NC + CI(s, i, j)]
#define PX(s, T, d, i, j) (T)->upx[(d) * (s)->NFX + FXI(s, i, j)]
#define MX(s, T, d, i, j) (T)->umx[(d) * (s)->NFX + FXI(s, i, j)]
#define PY(s, T, d, i, j) (T)->upy[(d) * (s)->NFY + FYI(s, i, j)]
#define MY(s, T, d, i, j) (T)->umy[(d) * (s)->NFY + FYI(s, i, j)]
#define B0V(s, T, i, j) (T)->b0[VI(s, i, j)]
#define BTV(s, T, i, j) (T)->bt[VI(s, i, j)]

/* Gather the 13-vector at a cell / face (for closures that take uvect). */
static void cell_vec(ko_sim* s, tdata* T, int i, int j, double* v) {
    for (int d = 0; d < NF; ++d) v[d] = UF(s, T, d, i, j);
}

static double gam_cell(ko_sim* s, tdata* T, int i, int j) {
    return gamma_of(&s->cfg, UF(s, T, DBDX, i, j), UF(s, T, DBDY, i, j));
}

/* ------------------------------------------------------------------ */
/* Topographic data (src/MorphodynamicRHS.f90:308-578, src/dem.f90:360-393) */
/* ------------------------------------------------------------------ */

/* ComputeCellCentredTopographicData_tileID (src/MorphodynamicRHS.f90:308-360);
 * on a 1D grid the two-vertex forms of the lower vertex row (:348-357), dbdy
 * left as it is. */
static void cell_topo(ko_sim* s, tdata* T) {
    int nx = s->nx, ny = s->ny;
    if (s->oned) {
        for (int i = 0; i < nx; ++i) {
            UF(s, T, B0, i, 0) = 0.5 * (B0V(s, T, i, 0) + B0V(s, T, i + 1, 0));
            UF(s, T, BT, i, 0) = 0.5 * (BTV(s, T, i, 0) + BTV(s, T, i + 1, 0));
            double a[4] = {B0V(s, T, i + 1, 0), BTV(s, T, i + 1, 0), -B0V(s, T, i, 0), -BTV(s, T, i, 0)};
            UF(s, T, DBDX, i, 0) = s->dxr * kahan_sum(a, 4);
        }
        return;
    }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            double a[8];
            a[0] = B0V(s, T, i, j); a[1] = B0V(s, T, i + 1, j);
            a[2] = B0V(s, T, i, j + 1); a[3] = B0V(s, T, i + 1, j + 1);
            double b0c = 0.25 * kahan_sum(a, 4);
            a[0] = BTV(s, T, i, j); a[1] = BTV(s, T, i + 1, j);
            a[2] = BTV(s, T, i, j + 1); a[3] = BTV(s, T, i + 1, j + 1);
            double btc = 0.25 * kahan_sum(a, 4);
            UF(s, T, B0, i, j) = b0c;
            UF(s, T, BT, i, j) = btc;
            a[0] = B0V(s, T, i + 1, j); a[1] = BTV(s, T, i + 1, j);
            a[2] = -B0V(s, T, i, j); a[3] = -BTV(s, T, i, j);
            a[4] = B0V(s, T, i + 1, j + 1); a[5] = BTV(s, T, i + 1, j + 1);
            a[6] = -B0V(s, T, i, j + 1); a[7] = -BTV(s, T, i, j + 1);
            UF(s, T, DBDX, i, j) = 0.5 * s->dxr * kahan_sum(a, 8);
            a[0] = B0V(s, T, i, j + 1); a[1] = BTV(s, T, i, j + 1);
            a[2] = -B0V(s, T, i, j); a[3] = -BTV(s, T, i, j);
            a[4] = B0V(s, T, i + 1, j + 1); a[5] = BTV(s, T, i + 1, j + 1);
            a[6] = -B0V(s, T, i + 1, j); a[7] = -BTV(s, T, i + 1, j);
            UF(s, T, DBDY, i, j) = 0.5 * s->dyr * kahan_sum(a, 8);
        }
}

/* Face b0 of GetHeights (src/dem.f90:378-386): vertex b0 of the tile from the
 * global (already single-valued) vertex array. */
static void get_heights(ko_sim* s, int k, tdata* T) {
    int nx = s->nx, ny = s->ny, gi, gj;
    tile_coords(s, k, &gi, &gj);
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i)
            B0V(s, T, i, j) = s->gb0[(size_t)(gj * ny + j) * (s->NX + 1) + gi * nx + i];
    if (s->oned) {   /* the lower vertex row (src/dem.f90:387-390) */
        for (int i = 0; i <= nx; ++i) {
            PX(s, T, B0, i, 0) = B0V(s, T, i, 0);
            MX(s, T, B0, i, 0) = B0V(s, T, i, 0);
        }
        return;
    }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            double v = 0.5 * (B0V(s, T, i, j) + B0V(s, T, i, j + 1));
            PX(s, T, B0, i, j) = v;
            MX(s, T, B0, i, j) = v;
        }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j <= ny; ++j) {
            double v = 0.5 * (B0V(s, T, i, j) + B0V(s, T, i + 1, j));
            PY(s, T, B0, i, j) = v;
            MY(s, T, B0, i, j) = v;
        }
}

/* ComputeInterfacialTopographicData (src/MorphodynamicRHS.f90:419-578); on a
 * 1D grid (:545-577) the face bt is the vertex's and dbdx a two-vertex
 * central difference of the lower row. */
static void face_topo(ko_sim* s, tdata* C, int k) {
    int nx = s->nx, ny = s->ny;
    tdata* T = &C[k];
    tdata* Wt = &C[s->meta[k].nb[0]];
    tdata* Et = &C[s->meta[k].nb[1]];
    double a[8];
    if (s->oned) {
        for (int i = 0; i <= nx; ++i) {
            tdata* L = (i == 0) ? Wt : T;          /* vertex i - 1 */
            int il = (i == 0) ? nx - 1 : i - 1;
            tdata* R = (i == nx) ? Et : T;         /* vertex i + 1 */
            int ir = (i == nx) ? 1 : i + 1;
            double bt = BTV(s, T, i, 0);
            a[0] = B0V(s, R, ir, 0); a[1] = BTV(s, R, ir, 0);
            a[2] = -B0V(s, L, il, 0); a[3] = -BTV(s, L, il, 0);
            double dbdx = 0.5 * s->dxr * kahan_sum(a, 4);
            PX(s, T, BT, i, 0) = bt; MX(s, T, BT, i, 0) = bt;
            PX(s, T, DBDX, i, 0) = dbdx; MX(s, T, DBDX, i, 0) = dbdx;
        }
        return;
    }
    tdata* Nt = &C[s->meta[k].nb[2]];
    tdata* St = &C[s->meta[k].nb[3]];
    for (int i = 1; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            a[0] = B0V(s, T, i + 1, j); a[1] = BTV(s, T, i + 1, j);
            a[2] = B0V(s, T, i + 1, j + 1); a[3] = BTV(s, T, i + 1, j + 1);
            a[4] = -B0V(s, T, i - 1, j); a[5] = -BTV(s, T, i - 1, j);
            a[6] = -B0V(s, T, i - 1, j + 1); a[7] = -BTV(s, T, i - 1, j + 1);
            double v = 0.25 * s->dxr * kahan_sum(a, 8);
            MX(s, T, DBDX, i, j) = v; PX(s, T, DBDX, i, j) = v;
        }
    for (int j = 0; j < ny; ++j) {
        a[0] = B0V(s, T, 1, j); a[1] = BTV(s, T, 1, j);
        a[2] = B0V(s, T, 1, j + 1); a[3] = BTV(s, T, 1, j + 1);
        a[4] = -B0V(s, Wt, nx - 1, j); a[5] = -BTV(s, Wt, nx - 1, j);
        a[6] = -B0V(s, Wt, nx - 1, j + 1); a[7] = -BTV(s, Wt, nx - 1, j + 1);
        double v = 0.25 * s->dxr * kahan_sum(a, 8);
        MX(s, T, DBDX, 0, j) = v; PX(s, T, DBDX, 0, j) = v;
    }
    for (int j = 0; j < ny; ++j) {
        a[0] = B0V(s, Et, 1, j); a[1] = BTV(s, Et, 1, j);
        a[2] = B0V(s, Et, 1, j + 1); a[3] = BTV(s, Et, 1, j + 1);
        a[4] = -B0V(s, T, nx - 1, j); a[5] = -BTV(s, T, nx - 1, j);
        a[6] = -B0V(s, T, nx - 1, j + 1); a[7] = -BTV(s, T, nx - 1, j + 1);
        double v = 0.25 * s->dxr * kahan_sum(a, 8);
        MX(s, T, DBDX, nx, j) = v; PX(s, T, DBDX, nx, j) = v;
    }
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j < ny; ++j) {
            double bt = 0.5 * (BTV(s, T, i, j) + BTV(s, T, i, j + 1));
            a[0] = B0V(s, T, i, j + 1); a[1] = BTV(s, T, i, j + 1);
            a[2] = -B0V(s, T, i, j); a[3] = -BTV(s, T, i, j);
            double dbdy = s->dyr * kahan_sum(a, 4);
            PX(s, T, BT, i, j) = bt; MX(s, T, BT, i, j) = bt;
            PX(s, T, DBDY, i, j) = dbdy; MX(s, T, DBDY, i, j) = dbdy;
        }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j <= ny; ++j) {
            double bt = 0.5 * (BTV(s, T, i, j) + BTV(s, T, i + 1, j));
            a[0] = B0V(s, T, i + 1, j); a[1] = BTV(s, T, i + 1, j);
            a[2] = -B0V(s, T, i, j); a[3] = -BTV(s, T, i, j);
            double dbdx = s->dxr * kahan_sum(a, 4);
            PY(s, T, BT, i, j) = bt; MY(s, T, BT, i, j) = bt;
            PY(s, T, DBDX, i, j) = dbdx; MY(s, T, DBDX, i, j) = dbdx;
        }
    for (int i = 0; i < nx; ++i)
        for (int j = 1; j < ny; ++j) {
            a[0] = B0V(s, T, i + 1, j + 1); a[1] = BTV(s, T, i + 1, j + 1);
            a[2] = B0V(s, T, i, j + 1); a[3] = BTV(s, T, i, j + 1);
            a[4] = -B0V(s, T, i + 1, j - 1); a[5] = -BTV(s, T, i + 1, j - 1);
            a[6] = -B0V(s, T, i, j - 1); a[7] = -BTV(s, T, i, j - 1);
            double v = 0.25 * s->dyr * kahan_sum(a, 8);
            PY(s, T, DBDY, i, j) = v; MY(s, T, DBDY, i, j) = v;
        }
    for (int i = 0; i < nx; ++i) {
        a[0] = B0V(s, T, i + 1, 1); a[1] = BTV(s, T, i + 1, 1);
        a[2] = B0V(s, T, i, 1); a[3] = BTV(s, T, i, 1);
        a[4] = -B0V(s, St, i + 1, ny - 1); a[5] = -BTV(s, St, i + 1, ny - 1);
        a[6] = -B0V(s, St, i, ny - 1); a[7] = -BTV(s, St, i, ny - 1);
        double v = 0.25 * s->dyr * kahan_sum(a, 8);
        PY(s, T, DBDY, i, 0) = v; MY(s, T, DBDY, i, 0) = v;
    }
    for (int i = 0; i < nx; ++i) {
        a[0] = B0V(s, Nt, i + 1, 1); a[1] = BTV(s, Nt, i + 1, 1);
        a[2] = B0V(s, Nt, i, 1); a[3] = BTV(s, Nt, i, 1);
        a[4] = -B0V(s, T, i + 1, ny - 1); a[5] = -BTV(s, T, i + 1, ny - 1);
        a[6] = -B0V(s, T, i, ny - 1); a[7] = -BTV(s, T, i, ny - 1);
        double v = 0.25 * s->dyr * kahan_sum(a, 8);
        PY(s, T, DBDY, i, ny) = v; MY(s, T, DBDY, i, ny) = v;
    }
}

/* ------------------------------------------------------------------ */
/* Hydraulic RHS (src/HydraulicRHS.f90)                               */
/* ------------------------------------------------------------------ */

/* CalculateLimitedDerivsBoundary (src/HydraulicRHS.f90:326-386), 2D. */
static void lim_bdry(ko_sim* s, tdata* C, int d, int i, int j, int k, char bid) {
    const kst_config* c = &s->cfg;
    tdata* T = &C[k];
    int nx = s->nx, ny = s->ny;
    switch (bid) {
    case 'N': {
        tdata* Nt = &C[s->meta[k].nb[2]];
        double uN = UF(s, Nt, d, i, 0);
        LY(s, T, d, i, j) = s->dyr * limiter(c, uN - UF(s, T, d, i, j), UF(s, T, d, i, j) - UF(s, T, d, i, j - 1));
        break;
    }
    case 'S': {
        tdata* St = &C[s->meta[k].nb[3]];
        double uS = UF(s, St, d, i, ny - 1);
        LY(s, T, d, i, j) = s->dyr * limiter(c, UF(s, T, d, i, j + 1) - UF(s, T, d, i, j), UF(s, T, d, i, j) - uS);
        break;
    }
    case 'E': {
        tdata* Et = &C[s->meta[k].nb[1]];
        double uE = UF(s, Et, d, 0, j);
        LX(s, T, d, i, j) = s->dxr * limiter(c, uE - UF(s, T, d, i, j), UF(s, T, d, i, j) - UF(s, T, d, i - 1, j));
        break;
    }
    default: {
        tdata* Wt = &C[s->meta[k].nb[0]];
        double uW = UF(s, Wt, d, nx - 1, j);
        LX(s, T, d, i, j) = s->dxr * limiter(c, UF(s, T, d, i + 1, j) - UF(s, T, d, i, j), UF(s, T, d, i, j) - uW);
        break;
    }
    }
}

static int tile_on(ko_sim* s, int k) { return in_bounds(s, k) && s->meta[k].on; }

/* CalculateLimitedDerivs (src/HydraulicRHS.f90:181-321), 2D. */
static void limited_derivs(ko_sim* s, tdata* C, const int* vars, int nv, int k) {
    const kst_config* c = &s->cfg;
    int nx = s->nx, ny = s->ny;
    tdata* T = &C[k];
    for (int j = 0; j < ny; ++j)
        for (int i = 1; i < nx - 1; ++i)
            for (int q = 0; q < nv; ++q) {
                int d = vars[q];
                LX(s, T, d, i, j) = s->dxr * limiter(c, UF(s, T, d, i + 1, j) - UF(s, T, d, i, j),
                                                     UF(s, T, d, i, j) - UF(s, T, d, i - 1, j));
            }
    for (int j = 1; j < ny - 1; ++j)
        for (int i = 0; i < nx; ++i)
            for (int q = 0; q < nv; ++q) {
                int d = vars[q];
                LY(s, T, d, i, j) = s->dyr * limiter(c, UF(s, T, d, i, j + 1) - UF(s, T, d, i, j),
                                                     UF(s, T, d, i, j) - UF(s, T, d, i, j - 1));
            }
    int nb;
    if (!s->oned) {   /* a 1D grid has no y slopes (they stay zero) */
        for (int i = 0; i < nx; ++i)
            for (int q = 0; q < nv; ++q) lim_bdry(s, C, vars[q], i, 0, k, 'S');
        nb = s->meta[k].nb[3];
        if (tile_on(s, nb))
            for (int i = 0; i < nx; ++i)
                for (int q = 0; q < nv; ++q) lim_bdry(s, C, vars[q], i, ny - 1, nb, 'N');
        for (int i = 0; i < nx; ++i)
            for (int q = 0; q < nv; ++q) lim_bdry(s, C, vars[q], i, ny - 1, k, 'N');
        nb = s->meta[k].nb[2];
        if (tile_on(s, nb))
            for (int i = 0; i < nx; ++i)
                for (int q = 0; q < nv; ++q) lim_bdry(s, C, vars[q], i, 0, nb, 'S');
    }
    for (int j = 0; j < ny; ++j)
        for (int q = 0; q < nv; ++q) lim_bdry(s, C, vars[q], 0, j, k, 'W');
    nb = s->meta[k].nb[0];
    if (tile_on(s, nb))
        for (int j = 0; j < ny; ++j)
            for (int q = 0; q < nv; ++q) lim_bdry(s, C, vars[q], nx - 1, j, nb, 'E');
    for (int j = 0; j < ny; ++j)
        for (int q = 0; q < nv; ++q) lim_bdry(s, C, vars[q], nx - 1, j, k, 'E');
    nb = s->meta[k].nb[1];
    if (tile_on(s, nb))
        for (int j = 0; j < ny; ++j)
            for (int q = 0; q < nv; ++q) lim_bdry(s, C, vars[q], 0, j, nb, 'W');
}

/* Reconstruct (src/HydraulicRHS.f90:397-552); on a 1D grid only the x faces,
 * and their rhoHnv keeps its reconstruction (:536-551 are 2D only). */
static void reconstruct(ko_sim* s, tdata* C, int desing, int k) {
    const kst_config* c = &s->cfg;
    int nx = s->nx, ny = s->ny;
    tdata* T = &C[k];
    tdata* Wt = &C[s->meta[k].nb[0]];
    tdata* Et = &C[s->meta[k].nb[1]];
    tdata* Nt = &C[s->meta[k].nb[2]];
    tdata* St = &C[s->meta[k].nb[3]];
    const int oned = s->oned;
    const int* dims = desing ? IDESING : IFLUX;
    int nd = desing ? 5 : 4;
    double hdx = 0.5, dx = c->dx, dy = c->dy;
    for (int q = 0; q < nd; ++q) {
        int d = dims[q];
        for (int j = 0; j < ny; ++j) {
            PX(s, T, d, 0, j) = UF(s, T, d, 0, j) - LX(s, T, d, 0, j) * hdx * dx;
            MX(s, T, d, 0, j) = UF(s, Wt, d, nx - 1, j) + LX(s, Wt, d, nx - 1, j) * hdx * dx;
            for (int i = 1; i < nx; ++i) {
                PX(s, T, d, i, j) = UF(s, T, d, i, j) - LX(s, T, d, i, j) * hdx * dx;
                MX(s, T, d, i, j) = UF(s, T, d, i - 1, j) + LX(s, T, d, i - 1, j) * hdx * dx;
            }
            PX(s, T, d, nx, j) = UF(s, Et, d, 0, j) - LX(s, Et, d, 0, j) * hdx * dx;
            MX(s, T, d, nx, j) = UF(s, T, d, nx - 1, j) + LX(s, T, d, nx - 1, j) * hdx * dx;
        }
        if (oned) continue;
        for (int i = 0; i < nx; ++i) {
            PY(s, T, d, i, 0) = UF(s, T, d, i, 0) - LY(s, T, d, i, 0) * hdx * dy;
            MY(s, T, d, i, 0) = UF(s, St, d, i, ny - 1) + LY(s, St, d, i, ny - 1) * hdx * dy;
            for (int j = 1; j < ny; ++j) {
                PY(s, T, d, i, j) = UF(s, T, d, i, j) - LY(s, T, d, i, j) * hdx * dy;
                MY(s, T, d, i, j) = UF(s, T, d, i, j - 1) + LY(s, T, d, i, j - 1) * hdx * dy;
            }
            PY(s, T, d, i, ny) = UF(s, Nt, d, i, 0) - LY(s, Nt, d, i, 0) * hdx * dy;
            MY(s, T, d, i, ny) = UF(s, T, d, i, ny - 1) + LY(s, T, d, i, ny - 1) * hdx * dy;
        }
    }
    if (!desing) return;
    for (int i = 0; i <= nx; ++i)
        for (int j = 0; j < ny; ++j) {
            double gam = gamma_of(c, PX(s, T, DBDX, i, j), PX(s, T, DBDY, i, j));
            PX(s, T, HN, i, j) = compute_hn(PX(s, T, W, i, j), PX(s, T, B0, i, j), PX(s, T, BT, i, j), gam);
            gam = gamma_of(c, MX(s, T, DBDX, i, j), MX(s, T, DBDY, i, j));
            MX(s, T, HN, i, j) = compute_hn(MX(s, T, W, i, j), MX(s, T, B0, i, j), MX(s, T, BT, i, j), gam);
        }
    if (oned) {
        for (int f = 0; f < s->NFX; ++f) {
            T->upx[RHU * s->NFX + f] = T->upx[RHO * s->NFX + f] * T->upx[HN * s->NFX + f] * T->upx[UU * s->NFX + f];
            T->umx[RHU * s->NFX + f] = T->umx[RHO * s->NFX + f] * T->umx[HN * s->NFX + f] * T->umx[UU * s->NFX + f];
        }
        return;
    }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j <= ny; ++j) {
            double gam = gamma_of(c, PY(s, T, DBDX, i, j), PY(s, T, DBDY, i, j));
            PY(s, T, HN, i, j) = compute_hn(PY(s, T, W, i, j), PY(s, T, B0, i, j), PY(s, T, BT, i, j), gam);
            gam = gamma_of(c, MY(s, T, DBDX, i, j), MY(s, T, DBDY, i, j));
            MY(s, T, HN, i, j) = compute_hn(MY(s, T, W, i, j), MY(s, T, B0, i, j), MY(s, T, BT, i, j), gam);
        }
    for (int f = 0; f < s->NFX; ++f) {
        T->upx[RHU * s->NFX + f] = T->upx[RHO * s->NFX + f] * T->upx[HN * s->NFX + f] * T->upx[UU * s->NFX + f];
        T->umx[RHU * s->NFX + f] = T->umx[RHO * s->NFX + f] * T->umx[HN * s->NFX + f] * T->umx[UU * s->NFX + f];
    }
    for (int f = 0; f < s->NFY; ++f) {
        T->upy[RHU * s->NFY + f] = T->upy[RHO * s->NFY + f] * T->upy[HN * s->NFY + f] * T->upy[UU * s->NFY + f];
        T->umy[RHU * s->NFY + f] = T->umy[RHO * s->NFY + f] * T->umy[HN * s->NFY + f] * T->umy[UU * s->NFY + f];
    }
    for (int f = 0; f < s->NFX; ++f) {
        T->upx[RHV * s->NFX + f] = T->upx[RHO * s->NFX + f] * T->upx[HN * s->NFX + f] * T->upx[VV * s->NFX + f];
        T->umx[RHV * s->NFX + f] = T->umx[RHO * s->NFX + f] * T->umx[HN * s->NFX + f] * T->umx[VV * s->NFX + f];
    }
    for (int f = 0; f < s->NFY; ++f) {
        T->upy[RHV * s->NFY + f] = T->upy[RHO * s->NFY + f] * T->upy[HN * s->NFY + f] * T->upy[VV * s->NFY + f];
        T->umy[RHV * s->NFY + f] = T->umy[RHO * s->NFY + f] * T->umy[HN * s->NFY + f] * T->umy[VV * s->NFY + f];
    }
}

/* InterpolateB (src/HydraulicRHS.f90:741-753). */
static double interp_b(double b00, double b01, double bt0, double bt1) {
    double b = bt0 + bt1;
    b = b + b00;
    b = b + b01;
    return b * 0.5;
}

/* The face bed CorrectSlopes compares with: InterpolateB of the face's two
 * vertices, or on a 1D grid b0 + bt of its lower vertex
 * (src/HydraulicRHS.f90:582-595). */
static double face_bed_x(ko_sim* s, tdata* T, int i, int j) {
    if (s->oned) return B0V(s, T, i, 0) + BTV(s, T, i, 0);
    return interp_b(B0V(s, T, i, j), B0V(s, T, i, j + 1), BTV(s, T, i, j), BTV(s, T, i, j + 1));
}

/* CorrectSlopes (src/HydraulicRHS.f90:560-736); a 1D grid has x faces only. */
static void correct_slopes(ko_sim* s, tdata* C, int k) {
    int nx = s->nx, ny = s->ny;
    tdata* T = &C[k];
    for (int j = 0; j < ny; ++j) {
        int pv = s->meta[k].nb[0];
        tdata* P = &C[pv];
        double bl = face_bed_x(s, P, nx - 1, j);
        double br = face_bed_x(s, P, nx, j);
        if ((MX(s, P, W, nx, j) < br) || (PX(s, P, W, nx - 1, j) < bl)) {
            MX(s, T, W, 0, j) = UF(s, P, W, nx - 1, j) + 0.5 * (br - bl);
        } else if (MX(s, T, W, 0, j) != MX(s, P, W, nx, j) && !s->meta[pv].ghost) {
            MX(s, T, W, 0, j) = MX(s, P, W, nx, j);
        }
        if (MX(s, T, HNPSI, 0, j) < 0.0) MX(s, T, HNPSI, 0, j) = UF(s, P, HNPSI, nx - 1, j);

        for (int i = 0; i < nx; ++i) {
            bl = face_bed_x(s, T, i, j);
            br = face_bed_x(s, T, i + 1, j);
            if ((MX(s, T, W, i + 1, j) < br) || (PX(s, T, W, i, j) < bl)) {
                MX(s, T, W, i + 1, j) = UF(s, T, W, i, j) + 0.5 * (br - bl);
                PX(s, T, W, i, j) = UF(s, T, W, i, j) + 0.5 * (bl - br);
            }
            if ((MX(s, T, HNPSI, i + 1, j) < 0.0) || (PX(s, T, HNPSI, i, j) < 0.0)) {
                MX(s, T, HNPSI, i + 1, j) = UF(s, T, HNPSI, i, j);
                PX(s, T, HNPSI, i, j) = UF(s, T, HNPSI, i, j);
            }
        }

        int nxt = s->meta[k].nb[1];
        tdata* N = &C[nxt];
        bl = face_bed_x(s, N, 0, j);
        br = face_bed_x(s, N, 1, j);
        if ((MX(s, N, W, 1, j) < br) || (PX(s, N, W, 0, j) < bl)) {
            PX(s, T, W, nx, j) = UF(s, N, W, 0, j) + 0.5 * (bl - br);
        } else if (PX(s, T, W, nx, j) != PX(s, N, W, 0, j) && !s->meta[nxt].ghost) {
            PX(s, T, W, nx, j) = PX(s, N, W, 0, j);
        }
        if (PX(s, T, HNPSI, nx, j) < 0.0) PX(s, T, HNPSI, nx, j) = UF(s, N, HNPSI, 0, j);
    }
    if (s->oned) return;
    for (int i = 0; i < nx; ++i) {
        int pv = s->meta[k].nb[3];
        tdata* P = &C[pv];
        double bb = interp_b(B0V(s, P, i, ny - 1), B0V(s, P, i + 1, ny - 1), BTV(s, P, i, ny - 1), BTV(s, P, i + 1, ny - 1));
        double bt = interp_b(B0V(s, P, i, ny), B0V(s, P, i + 1, ny), BTV(s, P, i, ny), BTV(s, P, i + 1, ny));
        if ((MY(s, P, W, i, ny) < bt) || (PY(s, P, W, i, ny - 1) < bb)) {
            MY(s, T, W, i, 0) = UF(s, P, W, i, ny - 1) + 0.5 * (bt - bb);
        } else if (MY(s, T, W, i, 0) != MY(s, P, W, i, ny) && !s->meta[pv].ghost) {
            MY(s, T, W, i, 0) = MY(s, P, W, i, ny);
        }
        if (MY(s, T, HNPSI, i, 0) < 0.0) MY(s, T, HNPSI, i, 0) = UF(s, P, HNPSI, i, ny - 1);

        for (int j = 0; j < ny; ++j) {
            bb = interp_b(B0V(s, T, i, j), B0V(s, T, i + 1, j), BTV(s, T, i, j), BTV(s, T, i + 1, j));
            bt = interp_b(B0V(s, T, i, j + 1), B0V(s, T, i + 1, j + 1), BTV(s, T, i, j + 1), BTV(s, T, i + 1, j + 1));
            if ((MY(s, T, W, i, j + 1) < bt) || (PY(s, T, W, i, j) < bb)) {
                MY(s, T, W, i, j + 1) = UF(s, T, W, i, j) + 0.5 * (bt - bb);
                PY(s, T, W, i, j) = UF(s, T, W, i, j) + 0.5 * (bb - bt);
            }
            if ((MY(s, T, HNPSI, i, j + 1) < 0.0) || (PY(s, T, HNPSI, i, j) < 0.0)) {
                MY(s, T, HNPSI, i, j + 1) = UF(s, T, HNPSI, i, j);
                PY(s, T, HNPSI, i, j) = UF(s, T, HNPSI, i, j);
            }
        }

        int nxt = s->meta[k].nb[2];
        tdata* N = &C[nxt];
        bb = interp_b(B0V(s, N, i, 0), B0V(s, N, i + 1, 0), BTV(s, N, i, 0), BTV(s, N, i + 1, 0));
        bt = interp_b(B0V(s, N, i, 1), B0V(s, N, i + 1, 1), BTV(s, N, i, 1), BTV(s, N, i + 1, 1));
        if ((MY(s, N, W, i, 1) < bt) || (PY(s, N, W, i, 0) < bb)) {
            PY(s, T, W, i, ny) = UF(s, N, W, i, 0) + 0.5 * (bb - bt);
        } else if (PY(s, T, W, i, ny) != PY(s, N, W, i, 0) && !s->meta[nxt].ghost) {
            PY(s, T, W, i, ny) = PY(s, N, W, i, 0);
        }
        if (PY(s, T, HNPSI, i, ny) < 0.0) PY(s, T, HNPSI, i, ny) = UF(s, N, HNPSI, i, 0);
    }
}

/* ComputeDesingularisedVariables (src/HydraulicRHS.f90:762-878). */
static void desingularise(ko_sim* s, tdata* C, int k, int velocities) {
    const kst_config* c = &s->cfg;
    tdata* T = &C[k];
    double eps = c->height_threshold;
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j) {
            double rhu = UF(s, T, RHU, i, j);
            double hnpsi = UF(s, T, HNPSI, i, j);
            double gam = gam_cell(s, T, i, j);
            double hn = compute_hn(UF(s, T, W, i, j), UF(s, T, B0, i, j), UF(s, T, BT, i, j), gam);
            if (hn < 0.0) hn = 0.0;
            if (hnpsi < 0.0) hnpsi = 0.0;
            double den = hn * hn + fmax_(hn * hn, eps * eps);
            double psi = fmin_(2.0 * hn * hnpsi / den, c->max_pack);
            double rho = density(c, psi);
            UF(s, T, HN, i, j) = hn;
            UF(s, T, PSI, i, j) = psi;
            UF(s, T, RHO, i, j) = rho;
            if (velocities) {
                UF(s, T, UU, i, j) = 2.0 * hn * rhu / (hn * hn + fmax_(hn * hn, eps * eps)) / rho;
                if (!s->oned) {   /* a 1D grid keeps v (src/HydraulicRHS.f90:870-874) */
                    double rhv = UF(s, T, RHV, i, j);
                    UF(s, T, VV, i, j) = 2.0 * hn * rhv / (hn * hn + fmax_(hn * hn, eps * eps)) / rho;
                }
            }
        }
}

/* Wave speeds (src/Equations.f90:249-381). */
static double wave_root(const kst_config* c, double hn, double bperp, double gam) {
    if (hn <= 0.0) hn = 0.0;
    if (c->geometric_factors) return sqrt(c->g * hn * (1.0 + bperp * bperp) / (gam * gam * gam));
    return sqrt(c->g * hn);
}

/* CalculateFluxes (src/HydraulicRHS.f90:896-1182) with the pointwise fluxes of
 * src/Equations.f90:53-245, 2D. Returns the tile's unit CFL step. */
static double calc_fluxes(ko_sim* s, tdata* C, int k) {
    const kst_config* c = &s->cfg;
    int nx = s->nx, ny = s->ny;
    tdata* T = &C[k];
    tdata* Wt = &C[s->meta[k].nb[0]];
    tdata* Et = &C[s->meta[k].nb[1]];
    tdata* Nt = &C[s->meta[k].nb[2]];
    tdata* St = &C[s->meta[k].nb[3]];
    double unit = HUGE_WP;
    double g = c->g, nu = c->eddy_viscosity;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            int f = FXI(s, i, j);
#define FP(d) T->upx[(d) * s->NFX + f]
#define FM(d) T->umx[(d) * s->NFX + f]
            double gP = gamma_of(c, FP(DBDX), FP(DBDY));
            double gM = gamma_of(c, FM(DBDX), FM(DBDY));
            double rP = wave_root(c, FP(HN), FP(DBDY), gP);
            double rM = wave_root(c, FM(HN), FM(DBDY), gM);
            double wsP = FP(UU) + rP, wsM = FM(UU) + rM;
            double aPos = (wsP > wsM) ? wsP : wsM;
            if (aPos < 0.0) aPos = 0.0;
            wsP = FP(UU) - rP; wsM = FM(UU) - rM;
            double aNeg = (wsP < wsM) ? wsP : wsM;
            if (aNeg > 0.0) aNeg = 0.0;
            if (aPos > EPS_WP) {
                double gc = (i == 0) ? gam_cell(s, Wt, nx - 1, j) : gam_cell(s, T, i - 1, j);
                double gr = fmin_(gc / gP, 1.0);
                unit = fmin_(gr * gr * c->dx / aPos, unit);
            }
            if (fabs(aNeg) > EPS_WP) {
                double gc = (i == nx) ? gam_cell(s, Et, 0, j) : gam_cell(s, T, i, j);
                double gr = fmin_(gc / gM, 1.0);
                unit = fmin_(gr * gr * c->dx / fabs(aNeg), unit);
            }
            /* XConvectionFlux */
            double cP[4], cM[4], hP[4], hM[4], pP[4], pM[4];
            cP[W] = FP(HN) * FP(UU) * gP; cP[HNPSI] = FP(HNPSI) * FP(UU) * gP;
            cP[RHU] = FP(RHU) * FP(UU); cP[RHV] = FP(RHV) * FP(UU);
            cM[W] = FM(HN) * FM(UU) * gM; cM[HNPSI] = FM(HNPSI) * FM(UU) * gM;
            cM[RHU] = FM(RHU) * FM(UU); cM[RHV] = FM(RHV) * FM(UU);
            /* XHydrostaticFlux */
            double hp = -FP(BT); hp = hp + (FP(W) - FP(B0));
            hP[W] = 0.0; hP[HNPSI] = 0.0; hP[RHU] = 0.5 * g * FP(RHO) * hp * hp; hP[RHV] = hP[RHU];
            hp = -FM(BT); hp = hp + (FM(W) - FM(B0));
            hM[W] = 0.0; hM[HNPSI] = 0.0; hM[RHU] = 0.5 * g * FM(RHO) * hp * hp; hM[RHV] = hM[RHU];
            /* XDiffusionFlux */
            const double* dP = (i == nx) ? &Et->ulx[CI(s, 0, j)] : &T->ulx[CI(s, i, j)];
            const double* dM = (i == 0) ? &Wt->ulx[CI(s, nx - 1, j)] : &T->ulx[CI(s, i - 1, j)];
            pP[W] = 0.0; pP[HNPSI] = 0.0; pM[W] = 0.0; pM[HNPSI] = 0.0;
            if (FP(HN) < 0.0) { pP[RHU] = 0.0; pP[RHV] = 0.0; }
            else { pP[RHU] = nu * FP(RHO) * FP(HN) * dP[UU * s->NC]; pP[RHV] = nu * FP(RHO) * FP(HN) * dP[VV * s->NC]; }
            if (FM(HN) < 0.0) { pM[RHU] = 0.0; pM[RHV] = 0.0; }
            else { pM[RHU] = nu * FM(RHO) * FM(HN) * dM[UU * s->NC]; pM[RHV] = nu * FM(RHO) * FM(HN) * dM[VV * s->NC]; }
            double dif = aPos - aNeg;
            for (int q = 0; q < 4; ++q) {
                int d = IFLUX[q];
                double* hX = &T->hx[d * s->NFX + f];
                double* gX = &T->gx[d * s->NFX + f];
                double* pX = &T->px[d * s->NFX + f];
                if (dif < 1e-10) { *hX = 0.0; *gX = 0.0; *pX = 0.0; }
                else if (d == W) {
                    double h = FP(HN) * gP - FM(HN) * gM;
                    h = h * aPos * aNeg;
                    h = h + (aPos * cM[q] - aNeg * cP[q]);
                    h = h / dif;
                    *hX = h; *gX = 0.0; *pX = 0.0;
                } else if (d == HNPSI) {
                    double h = FP(d) * gP - FM(d) * gM;
                    h = h * aPos * aNeg;
                    h = h + (aPos * cM[q] - aNeg * cP[q]);
                    h = h / dif;
                    *hX = h; *gX = 0.0; *pX = 0.0;
                } else {
                    double h = FP(d) - FM(d);
                    h = h * aPos * aNeg;
                    h = h + (aPos * cM[q] - aNeg * cP[q]);
                    h = h / dif;
                    *hX = h;
                    *gX = (aPos * hM[q] - aNeg * hP[q]) / dif;
                    *pX = 0.5 * (pP[q] + pM[q]);
                }
            }
#undef FP
#undef FM
        }
    if (s->oned) return unit;   /* no y fluxes on a 1D grid (src/HydraulicRHS.f90:1075) */
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int f = FYI(s, i, j);
#define FP(d) T->upy[(d) * s->NFY + f]
#define FM(d) T->umy[(d) * s->NFY + f]
            double gP = gamma_of(c, FP(DBDX), FP(DBDY));
            double gM = gamma_of(c, FM(DBDX), FM(DBDY));
            double rP = wave_root(c, FP(HN), FP(DBDX), gP);
            double rM = wave_root(c, FM(HN), FM(DBDX), gM);
            double wsP = FP(VV) + rP, wsM = FM(VV) + rM;
            double aPos = (wsP > wsM) ? wsP : wsM;
            if (aPos < 0.0) aPos = 0.0;
            wsP = FP(VV) - rP; wsM = FM(VV) - rM;
            double aNeg = (wsP < wsM) ? wsP : wsM;
            if (aNeg > 0.0) aNeg = 0.0;
            if (aPos > EPS_WP) {
                double gc = (j == 0) ? gam_cell(s, St, i, ny - 1) : gam_cell(s, T, i, j - 1);
                double gr = fmin_(gc / gP, 1.0);
                unit = fmin_(gr * gr * c->dy / aPos, unit);
            }
            if (fabs(aNeg) > EPS_WP) {
                double gc = (j == ny) ? gam_cell(s, Nt, i, 0) : gam_cell(s, T, i, j);
                double gr = fmin_(gc / gM, 1.0);
                unit = fmin_(gr * gr * c->dy / fabs(aNeg), unit);
            }
            double cP[4], cM[4], hP[4], hM[4], pP[4], pM[4];
            cP[W] = FP(HN) * FP(VV) * gP; cP[HNPSI] = FP(HNPSI) * FP(VV) * gP;
            cP[RHU] = FP(RHU) * FP(VV); cP[RHV] = FP(RHV) * FP(VV);
            cM[W] = FM(HN) * FM(VV) * gM; cM[HNPSI] = FM(HNPSI) * FM(VV) * gM;
            cM[RHU] = FM(RHU) * FM(VV); cM[RHV] = FM(RHV) * FM(VV);
            double hp = -FP(BT); hp = hp + (FP(W) - FP(B0));
            hP[W] = 0.0; hP[HNPSI] = 0.0; hP[RHU] = 0.5 * g * FP(RHO) * hp * hp; hP[RHV] = hP[RHU];
            hp = -FM(BT); hp = hp + (FM(W) - FM(B0));
            hM[W] = 0.0; hM[HNPSI] = 0.0; hM[RHU] = 0.5 * g * FM(RHO) * hp * hp; hM[RHV] = hM[RHU];
            const double* dP = (j == ny) ? &Nt->uly[CI(s, i, 0)] : &T->uly[CI(s, i, j)];
            const double* dM = (j == 0) ? &St->uly[CI(s, i, ny - 1)] : &T->uly[CI(s, i, j - 1)];
            pP[W] = 0.0; pP[HNPSI] = 0.0; pM[W] = 0.0; pM[HNPSI] = 0.0;
            if (FP(HN) < 0.0) { pP[RHU] = 0.0; pP[RHV] = 0.0; }
            else { pP[RHU] = nu * FP(RHO) * FP(HN) * dP[UU * s->NC]; pP[RHV] = nu * FP(RHO) * FP(HN) * dP[VV * s->NC]; }
            if (FM(HN) < 0.0) { pM[RHU] = 0.0; pM[RHV] = 0.0; }
            else { pM[RHU] = nu * FM(RHO) * FM(HN) * dM[UU * s->NC]; pM[RHV] = nu * FM(RHO) * FM(HN) * dM[VV * s->NC]; }
            double dif = aPos - aNeg;
            for (int q = 0; q < 4; ++q) {
                int d = IFLUX[q];
                double* hY = &T->hy[d * s->NFY + f];
                double* gY = &T->gy[d * s->NFY + f];
                double* pY = &T->py[d * s->NFY + f];
                if (dif < 1e-10) { *hY = 0.0; *gY = 0.0; *pY = 0.0; }
                else if (d == W) {
                    double h = FP(HN) * gP - FM(HN) * gM;
                    h = h * aPos * aNeg;
                    h = h + (aPos * cM[q] - aNeg * cP[q]);
                    h = h / dif;
                    *hY = h; *gY = 0.0; *pY = 0.0;
                } else if (d == HNPSI) {
                    double h = FP(d) * gP - FM(d) * gM;
                    h = h * aPos * aNeg;
                    h = h + (aPos * cM[q] - aNeg * cP[q]);
                    h = h / dif;
                    *hY = h; *gY = 0.0; *pY = 0.0;
                } else {
                    double h = FP(d) - FM(d);
                    h = h * aPos * aNeg;
                    h = h + (aPos * cM[q] - aNeg * cP[q]);
                    h = h / dif;
                    *hY = h;
                    *gY = (aPos * hM[q] - aNeg * hP[q]) / dif;
                    *pY = 0.5 * (pP[q] + pM[q]);
                }
            }
#undef FP
#undef FM
        }
    return unit;
}

/* ExplicitSourceTerms (src/Equations.f90:456-620). */
static void explicit_sources(ko_sim* s, double t, double x, double y, const double* uv, int has_src,
                             double* st) {
    const kst_config* c = &s->cfg;
    for (int q = 0; q < 4; ++q) st[q] = 0.0;
    double Qt = 0.0, psiQt = 0.0;
    if (has_src && c->n_sources >= 1) {
        double qsum = 0.0, psisum = 0.0;
        for (int J = 0; J < c->n_sources; ++J) {
            const kst_flux_source* fs = &s->src[J];
            double Qf = 0.0, psiQf = 0.0;
            if (((x - fs->x) * (x - fs->x) + (y - fs->y) * (y - fs->y)) < fs->radius * fs->radius) {
                int n = fs->n_series;
                if (n == 1) {
                    if (t < fs->time[0] || (t == fs->time[0] && s->t < t)) {
                        Qf = 0.0; psiQf = 0.0;
                    } else {
                        Qf = fs->flux[0];
                        psiQf = fs->psi[0] * Qf;
                    }
                } else {
                    if (t < fs->time[0] || t > fs->time[n - 1] || (t == fs->time[0] && s->t < t) ||
                        (t == fs->time[n - 1] && s->t == t)) {
                        Qf = 0.0; psiQf = 0.0;
                    } else {
                        for (int K = 1; K < n; ++K) {
                            if (t >= fs->time[K - 1] && t <= fs->time[K]) {
                                double Qa = fs->flux[K - 1], psia = fs->psi[K - 1], ta = fs->time[K - 1];
                                double Qb = fs->flux[K], psib = fs->psi[K], tb = fs->time[K];
                                Qf = Qa + (Qb - Qa) * (t - ta) / (tb - ta);
                                double psif = psia + (psib - psia) * (t - ta) / (tb - ta);
                                psiQf = psif * Qf;
                            }
                        }
                    }
                }
            }
            if (s->oned) {   /* per unit length on a 1D grid (src/Equations.f90:582-585) */
                Qf = Qf / (double)fs->num_cells / c->dx;
                psiQf = psiQf / (double)fs->num_cells / c->dx;
            } else {
                Qf = Qf / (double)fs->num_cells / c->dx / c->dy;
                psiQf = psiQf / (double)fs->num_cells / c->dx / c->dy;
            }
            qsum = qsum + Qf;
            psisum = psisum + psiQf;
        }
        Qt = Qt + qsum;
        psiQt = psiQt + psisum;
    }
    double gam = gamma_of(c, uv[DBDX], uv[DBDY]);
    st[W] = st[W] + Qt / (gam * gam);
    st[HNPSI] = st[HNPSI] + psiQt / gam;
    double hpg = -uv[BT];
    hpg = hpg + (uv[W] - uv[B0]);
    hpg = hpg / gam;
    st[RHU] = st[RHU] - c->g * uv[RHO] * hpg * uv[DBDX];
    st[RHV] = st[RHV] - c->g * uv[RHO] * hpg * uv[DBDY];
}

/* The implicit drag terms of a cell (ImplicitSourceTerms,
 * src/Equations.f90:627-658) and its explicit source terms. */
static void cell_sources(ko_sim* s, tdata* T, int k, int i, int j, double t, const double* stf) {
    const kst_config* c = &s->cfg;
    double uv[NF], ste[4];
    cell_vec(s, T, i, j, uv);
    explicit_sources(s, t, s->meta[k].x[i], s->meta[k].y[j], uv, s->meta[k].src, ste);
    for (int q = 0; q < 4; ++q) T->dte[q * s->NC + CI(s, i, j)] = stf[q] + ste[q];
    cstate cs = {uv[HN], uv[UU], uv[VV], uv[PSI], uv[BT], uv[DBDX], uv[DBDY]};
    double fric = drag(c, &cs);
    double sti[4] = {0, 0, 0, 0};
    if (uv[HN] > c->height_threshold) {
        double modu = sqrt(speed2(c, uv[UU], uv[VV], uv[DBDX], uv[DBDY]));
        if (modu > 1.0e-8) {
            double hr = 1.0 / uv[HN];
            sti[RHU] = -(fric * hr / modu);
            sti[RHV] = -(fric * hr / modu);
        }
    }
    for (int q = 0; q < 4; ++q) T->dti[q * s->NC + CI(s, i, j)] = sti[q];
}

/* ConstructHydraulicRHS on a 1D grid (src/HydraulicRHS.f90:1276-1302): x
 * fluxes only, the hydrostatic flux difference divided by gamma; the first
 * (plain-sum) expression of the momentum terms is overwritten by the Kahan
 * sum. */
static void construct_rhs_1d(ko_sim* s, tdata* C, int k, double t) {
    tdata* T = &C[k];
    for (int i = 0; i < s->nx; ++i) {
        double gam = gam_cell(s, T, i, 0);
        double stf[4];
        int fx0 = FXI(s, i, 0), fx1 = FXI(s, i + 1, 0);
        for (int q = 0; q < 4; ++q) {
            int d = IFLUX[q];
            const double* hX = &T->hx[d * s->NFX];
            const double* gX = &T->gx[d * s->NFX];
            const double* pX = &T->px[d * s->NFX];
            if (d == W) {
                stf[q] = (hX[fx0] - hX[fx1]) * s->dxr / (gam * gam);
            } else if (d == HNPSI) {
                stf[q] = (hX[fx0] - hX[fx1]) * s->dxr / gam;
            } else {
                double a[3];
                a[0] = hX[fx0] - hX[fx1];
                a[1] = (gX[fx0] - gX[fx1]) / gam;
                a[2] = pX[fx1] - pX[fx0];
                stf[q] = kahan_sum(a, 3) * s->dxr;
            }
        }
        cell_sources(s, T, k, i, 0, t, stf);
    }
}

/* ConstructHydraulicRHS (src/HydraulicRHS.f90:1190-1304) with
 * ImplicitSourceTerms (src/Equations.f90:627-658). */
static void construct_rhs(ko_sim* s, tdata* C, int k, double t) {
    const kst_config* c = &s->cfg;
    tdata* T = &C[k];
    int nx = s->nx, ny = s->ny;
    if (s->oned) {
        construct_rhs_1d(s, C, k, t);
        return;
    }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            double gam = gam_cell(s, T, i, j);
            double bx = UF(s, T, DBDX, i, j), by = UF(s, T, DBDY, i, j);
            double gxp[4] = {0, 0, 0, 0}, gyp[4] = {0, 0, 0, 0};
            if (c->geometric_factors) {
                gxp[RHU] = (1.0 + by * by) / gam;
                gxp[RHV] = -(bx * by / gam);
                gyp[RHU] = -(bx * by / gam);
                gyp[RHV] = (1.0 + bx * bx) / gam;
            } else {
                gxp[RHU] = 1.0; gxp[RHV] = 0.0; gyp[RHU] = 0.0; gyp[RHV] = 1.0;
            }
            double stf[4];
            int fx0 = FXI(s, i, j), fx1 = FXI(s, i + 1, j);
            int fy0 = FYI(s, i, j), fy1 = FYI(s, i, j + 1);
            for (int q = 0; q < 4; ++q) {
                int d = IFLUX[q];
                const double* hX = &T->hx[d * s->NFX];
                const double* gX = &T->gx[d * s->NFX];
                const double* pX = &T->px[d * s->NFX];
                const double* hY = &T->hy[d * s->NFY];
                const double* gY = &T->gy[d * s->NFY];
                const double* pY = &T->py[d * s->NFY];
                if (d == W) {
                    stf[q] = (hX[fx0] - hX[fx1]) * s->dxr / (gam * gam) + (hY[fy0] - hY[fy1]) * s->dyr / (gam * gam);
                } else if (d == HNPSI) {
                    stf[q] = (hX[fx0] - hX[fx1]) * s->dxr / gam + (hY[fy0] - hY[fy1]) * s->dyr / gam;
                } else {
                    double a[3];
                    a[0] = hX[fx0] - hX[fx1];
                    a[1] = (gX[fx0] - gX[fx1]) * gxp[q];
                    a[2] = pX[fx1] - pX[fx0];
                    stf[q] = kahan_sum(a, 3) * s->dxr;
                    a[0] = hY[fy0] - hY[fy1];
                    a[1] = (gY[fy0] - gY[fy1]) * gyp[q];
                    a[2] = pY[fy1] - pY[fy0];
                    stf[q] = stf[q] + kahan_sum(a, 3) * s->dyr;
                }
            }
            cell_sources(s, T, k, i, j, t, stf);
        }
}

/* ComputeAdvisedTimeStep (src/HydraulicRHS.f90:141-174). */
static double advised_step(ko_sim* s, int substep, double min_unit, int set_dt) {
    const kst_config* c = &s->cfg;
    double m = fmin_(c->cfl * min_unit, c->diffusive_time_scale);
    m = fmin_(m, c->max_dt);
    if (substep == 1) {
        double a = 0.9 * m;
        if (set_dt) s->dt = a;
        return a;
    }
    return m;
}

/* CalculateHydraulicRHS (src/HydraulicRHS.f90:64-136). */
static double hydraulic_rhs(ko_sim* s, tdata* C, double t, int substep) {
    int na = s->nactive;
    double min_unit = HUGE_WP;
    /* unitCFLTimeStep(:) = maxdt for every tile, active ones overwritten */
    if (na < s->ntiles) min_unit = s->cfg.max_dt;
    for (int q = 0; q < na; ++q) limited_derivs(s, C, IFLUX, 4, s->active[q]);
    for (int q = 0; q < na; ++q) reconstruct(s, C, 0, s->active[q]);
    for (int q = 0; q < na; ++q) correct_slopes(s, C, s->active[q]);
    for (int q = 0; q < na; ++q) desingularise(s, C, s->active[q], 1);
    for (int q = 0; q < na; ++q) limited_derivs(s, C, IDESING, 5, s->active[q]);
    for (int q = 0; q < na; ++q) reconstruct(s, C, 1, s->active[q]);
    for (int q = 0; q < na; ++q) {
        double u = calc_fluxes(s, C, s->active[q]);
        min_unit = fmin_(min_unit, u);
    }
    double adv = advised_step(s, substep, min_unit, 1);
    for (int q = 0; q < na; ++q) construct_rhs(s, C, s->active[q], t);
    return adv;
}

/* ------------------------------------------------------------------ */
/* Tile activation (src/UpdateTiles.f90, src/TimeStepper.f90:924-1150) */
/* ------------------------------------------------------------------ */

static int in_list(const int* l, int n, int k) {
    for (int i = 0; i < n; ++i)
        if (l[i] == k) return 1;
    return 0;
}

/* SetTileBoundaries (src/UpdateTiles.f90:484-512) and, with periodic
 * boundaries, SetPeriodicBoundaryConds (:521-566): a tile in the first or
 * last column gets the opposite column's tile as its west/east neighbour and
 * its diagonal tiles from there; a tile in the first or last row gets the
 * opposite row's as its south/north neighbour (its diagonals stay the
 * plain-arithmetic ones, which periodic ghosts never use). */
static void set_tile_boundaries(ko_sim* s, int k) {
    tmeta* m = &s->meta[k];
    int ntx = s->ntx, nty = s->nty;
    m->nb[0] = k - 1; m->nb[1] = k + 1; m->nb[2] = k + ntx; m->nb[3] = k - ntx;
    m->cr[0] = (k - ntx) - 1; m->cr[1] = (k - ntx) + 1; m->cr[2] = (k + ntx) - 1; m->cr[3] = (k + ntx) + 1;
    if (s->cfg.bc != KST_BC_PERIODIC) return;
    int gi, gj;
    tile_coords(s, k, &gi, &gj);
#define TID(i, j) ((i) + (j) * ntx)
    if (gi == 0) {
        m->nb[0] = TID(ntx - 1, gj);
        if (gj > 0 && gj < nty - 1) { m->cr[0] = TID(ntx - 1, gj - 1); m->cr[2] = TID(ntx - 1, gj + 1); }
        else if (gj == 0) { m->cr[0] = TID(ntx - 1, nty - 1); m->cr[2] = TID(ntx - 1, gj + 1); }
        else if (gj == nty - 1) { m->cr[0] = TID(ntx - 1, gj - 1); m->cr[2] = TID(ntx - 1, 0); }
    }
    if (gi == ntx - 1) {
        m->nb[1] = TID(0, gj);
        if (gj > 0 && gj < nty - 1) { m->cr[1] = TID(0, gj - 1); m->cr[3] = TID(0, gj + 1); }
        else if (gj == 0) { m->cr[1] = TID(0, nty - 1); m->cr[3] = TID(0, gj + 1); }
        else if (gj == nty - 1) { m->cr[1] = TID(0, gj - 1); m->cr[3] = TID(0, 0); }
    }
    if (gj == 0) m->nb[3] = TID(gi, nty - 1);
    if (gj == nty - 1) m->nb[2] = TID(gi, 0);
#undef TID
}

/* AllocateTopographicData (src/UpdateTiles.f90:271-337): coordinates via
 * GridToPhysical (src/Grid.f90:339-353). */
static void alloc_topo(ko_sim* s, int k) {
    tmeta* m = &s->meta[k];
    if (m->x) return;
    const kst_config* c = &s->cfg;
    int gi, gj;
    tile_coords(s, k, &gi, &gj);
    m->x = xcalloc(s->nx, 8); m->y = xcalloc(s->ny, 8);
    m->xv = xcalloc(s->nx + 1, 8); m->yv = xcalloc(s->ny + 1, 8);
    for (int ii = 0; ii < s->nx; ++ii)
        m->x[ii] = -0.5 * c->x_size + c->dx * (((gi + 1) - 1.0) * s->nx + ((ii + 1) - 0.5));
    for (int jj = 0; jj < s->ny; ++jj)
        m->y[jj] = -0.5 * c->y_size + c->dy * (((gj + 1) - 1.0) * s->ny + ((jj + 1) - 0.5));
    m->xmin = m->x[0]; m->xmax = m->x[s->nx - 1];
    m->ymin = m->y[0]; m->ymax = m->y[s->ny - 1];
}

static void alloc_maxima(ko_sim* s, int k) {
    tmeta* m = &s->meta[k];
    if (m->hnmax) return;
    int NC = s->NC;
    m->hnmax = xcalloc(2 * NC, 8); m->umax = xcalloc(2 * NC, 8); m->emax = xcalloc(2 * NC, 8);
    m->dmax = xcalloc(2 * NC, 8); m->psimax = xcalloc(2 * NC, 8);
    m->tfirst = xcalloc(NC, 8);
    for (int i = 0; i < NC; ++i) m->tfirst[i] = -1.0;
}

/* AllocateU (src/UpdateTiles.f90:222-269): zero data with rho = rhow. */
static void alloc_u(ko_sim* s, int k) {
    tdata* T = &s->cont[C_TC][k];
    if (T->alloc) return;
    tdata_alloc(s, T);
    for (int i = 0; i < s->NC; ++i) T->u[RHO * s->NC + i] = s->cfg.rhow;
    s->meta[k].touched = 1;
}

/* ReconstructwAtEdges (src/UpdateTiles.f90:669-750), 2D. */
static void reconstruct_w_at_edges(ko_sim* s, int k) {
    tdata* C = s->cont[C_TC];
    tdata* T = &C[k];
    int nx = s->nx, ny = s->ny;
    double dx = s->cfg.dx, dy = s->cfg.dy;
    int nb = s->meta[k].nb[0];
    if (in_bounds(s, nb) && s->meta[nb].on) {
        for (int j = 0; j < ny; ++j) lim_bdry(s, C, W, 0, j, k, 'W');
        for (int j = 0; j < ny; ++j) {
            PX(s, T, W, 0, j) = UF(s, T, W, 0, j) - LX(s, T, W, 0, j) * 0.5 * dx;
            MX(s, T, W, 1, j) = UF(s, T, W, 0, j) + LX(s, T, W, 0, j) * 0.5 * dx;
        }
    }
    nb = s->meta[k].nb[1];
    if (in_bounds(s, nb) && s->meta[nb].on) {
        for (int j = 0; j < ny; ++j) lim_bdry(s, C, W, nx - 1, j, k, 'E');
        for (int j = 0; j < ny; ++j) {
            PX(s, T, W, nx - 1, j) = UF(s, T, W, nx - 1, j) - LX(s, T, W, nx - 1, j) * 0.5 * dx;
            MX(s, T, W, nx, j) = UF(s, T, W, nx - 1, j) + LX(s, T, W, nx - 1, j) * 0.5 * dx;
        }
    }
    if (s->oned) return;   /* x edges only on a 1D grid (src/UpdateTiles.f90:718) */
    nb = s->meta[k].nb[3];
    if (in_bounds(s, nb) && s->meta[nb].on) {
        for (int i = 0; i < nx; ++i) lim_bdry(s, C, W, i, 0, k, 'S');
        for (int i = 0; i < nx; ++i) {
            PY(s, T, W, i, 0) = UF(s, T, W, i, 0) - LY(s, T, W, i, 0) * 0.5 * dy;
            MY(s, T, W, i, 1) = UF(s, T, W, i, 0) + LY(s, T, W, i, 0) * 0.5 * dy;
        }
    }
    nb = s->meta[k].nb[2];
    if (in_bounds(s, nb) && s->meta[nb].on) {
        for (int i = 0; i < nx; ++i) lim_bdry(s, C, W, i, ny - 1, k, 'N');
        for (int i = 0; i < nx; ++i) {
            PY(s, T, W, i, ny - 1) = UF(s, T, W, i, ny - 1) - LY(s, T, W, i, ny - 1) * 0.5 * dy;
            MY(s, T, W, i, ny) = UF(s, T, W, i, ny - 1) + LY(s, T, W, i, ny - 1) * 0.5 * dy;
        }
    }
}

/* SetDomainBoundaryData (src/UpdateTiles.f90:590-643) for a ghost tile on
 * the domain edge with Dirichlet boundaries: w from the boundary depth over
 * the cell's own bed, the momenta, solids and derived fields constant. */
static void set_dirichlet_tile_data(ko_sim* s, int k) {
    const kst_config* c = &s->cfg;
    tdata* T = &s->cont[C_TC][k];
    const double hn = c->bc_hn, uv = c->bc_u, vv = c->bc_v, ps = c->bc_psi;
    const double rho = c->rhow + (c->rhos - c->rhow) * ps;   /* Density */
    for (int i = 0; i < s->NC; ++i) {
        const double hp = hn / gamma_of(c, T->u[DBDX * s->NC + i], T->u[DBDY * s->NC + i]);
        double w = T->u[BT * s->NC + i] + hp;
        w = w + T->u[B0 * s->NC + i];
        T->u[W * s->NC + i] = w;
        T->u[RHU * s->NC + i] = rho * hn * uv;
        T->u[RHV * s->NC + i] = rho * hn * vv;
        T->u[HNPSI * s->NC + i] = hn * ps;
        T->u[HN * s->NC + i] = hn;
        T->u[UU * s->NC + i] = uv;
        T->u[VV * s->NC + i] = vv;
        T->u[PSI * s->NC + i] = ps;
        T->u[RHO * s->NC + i] = rho;
    }
}

/* SetDefaultTileData (src/UpdateTiles.f90:647-664). */
static void set_default_tile_data(ko_sim* s, int k) {
    tdata* T = &s->cont[C_TC][k];
    for (int d = 0; d < 9; ++d)
        for (int i = 0; i < s->NC; ++i) T->u[d * s->NC + i] = 0.0;
    for (int i = 0; i < s->NC; ++i) {
        T->u[RHO * s->NC + i] = s->cfg.rhow;
        T->u[W * s->NC + i] = T->u[B0 * s->NC + i];
    }
}

/* AddGhostTiles (src/UpdateTiles.f90:389-481): the four neighbours (W and E
 * on a 1D grid) and, but with periodic boundaries, the four diagonal tiles.  A new ghost gets
 * UpdateGhostTileBCs's data (:571-585): on the domain edge
 * SetDomainBoundaryData's -- none with periodic boundaries (AllocateU's zeros
 * and rho = rhow stay), the boundary data with Dirichlet ones, else the
 * default dry data -- and off the edge the default dry data. */
static int add_ghost_tiles(ko_sim* s, int k) {
    int newg[8], ng = 0;
    /* grid%dim*2 + 4*(grid%dim - 1) candidates, the corners skipped with
     * periodic boundaries: W, E on a 1D grid */
    const int nq = s->oned ? 2 : (s->cfg.bc == KST_BC_PERIODIC ? 4 : 8);
    for (int q = 0; q < nq; ++q) {
        int tile = (q < 4) ? s->meta[k].nb[q] : s->meta[k].cr[q - 4];
        if (!in_bounds(s, tile)) return fail(s, KST_ERR_HALT, "ghost tile outside the domain");
        int already = in_list(s->ghosts, s->nghost, tile);
        if (already) {
            reconstruct_w_at_edges(s, tile);
        } else if (!s->meta[tile].on && !in_list(newg, ng, tile)) {
            newg[ng++] = tile;
        }
    }
    if (ng == 0) return 0;
    for (int q = 0; q < ng; ++q) s->ghosts[s->nghost++] = newg[q];
    for (int q = 0; q < ng; ++q) {
        int tile = newg[q];
        alloc_u(s, tile);
        alloc_topo(s, tile);
        set_tile_boundaries(s, tile);
        get_heights(s, tile, &s->cont[C_TC][tile]);
        cell_topo(s, &s->cont[C_TC][tile]);
        if (!on_domain_edge(s, tile)) set_default_tile_data(s, tile);
        else if (s->cfg.bc == KST_BC_DIRICHLET) set_dirichlet_tile_data(s, tile);
        else if (s->cfg.bc != KST_BC_PERIODIC) set_default_tile_data(s, tile);
        reconstruct_w_at_edges(s, tile);
        s->meta[tile].ghost = 1;
    }
    /* EqualiseTopographicBoundaryData is a no-op on a single-valued b0. */
    return 0;
}

/* AddToActiveTiles (src/UpdateTiles.f90:81-117). */
static void add_to_active(ko_sim* s, int k) {
    if (in_list(s->active, s->nactive, k)) return;
    int p = s->nactive;
    while (p > 0 && s->active[p - 1] > k) { s->active[p] = s->active[p - 1]; --p; }
    s->active[p] = k;
    s->nactive++;
    for (int q = 0; q < s->nghost; ++q)
        if (s->ghosts[q] == k) {
            for (int r = q; r < s->nghost - 1; ++r) s->ghosts[r] = s->ghosts[r + 1];
            s->nghost--;
            s->meta[k].ghost = 0;
            break;
        }
}

/* ActivateTile (src/UpdateTiles.f90:342-384): the pre-pass only touches the
 * new tile's own work arrays, which every later RHS recomputes. */
static int activate_tile(ko_sim* s, int k) {
    tdata* T = &s->cont[C_TC][k];
    s->meta[k].ghost = 0;
    for (int i = 0; i < s->NC; ++i) {
        T->u[W * s->NC + i] = T->u[B0 * s->NC + i];
        T->u[BT * s->NC + i] = 0.0;
    }
    int rc = add_ghost_tiles(s, k);
    if (rc) return rc;
    face_topo(s, s->cont[C_TC], k);
    desingularise(s, s->cont[C_TC], k, 1);
    return 0;
}

/* AddTile (src/UpdateTiles.f90:56-78): a tile outside the domain, or on its
 * edge but with periodic boundaries, is fatal with 'halt' and silently not
 * added otherwise. */
static int add_tile(ko_sim* s, int k) {
    if (!in_bounds(s, k) || (on_domain_edge(s, k) && s->cfg.bc != KST_BC_PERIODIC)) {
        if (s->cfg.bc == KST_BC_HALT)
            return fail(s, KST_ERR_HALT, "Error: tried to add a tile outside the domain (tile %d)", k);
        return 0;
    }
    add_to_active(s, k);
    /* AllocateTile (src/UpdateTiles.f90:120-219) */
    s->meta[k].on = 1;
    alloc_maxima(s, k);
    alloc_u(s, k);
    set_tile_boundaries(s, k);
    alloc_topo(s, k);
    get_heights(s, k, &s->cont[C_TC][k]);
    cell_topo(s, &s->cont[C_TC][k]);
    s->meta[k].src = 0;
    return activate_tile(s, k);
}

/* Near{North,South,East,West}Boundary (src/TimeStepper.f90:946-1150). The
 * loops re-read the (sorted, growing) active list by index as the reference
 * does. */
static int check_near_boundaries(ko_sim* s) {
    const kst_config* c = &s->cfg;
    tdata* C = s->cont[C_TC];
    int nx = s->nx, ny = s->ny, tb = c->tile_buffer;
    double thr = c->height_threshold;
    int size, rc;
    /* North (and South) only on a 2D grid (src/TimeStepper.f90:935-938) */
    size = s->oned ? 0 : s->nactive;
    for (int tt = 0; tt < size; ++tt) {
        int k = s->active[tt], n = s->meta[k].nb[2];
        if (n > -1 && !(in_bounds(s, n) && s->meta[n].on)) {
            long yd = 0;
            for (int ii = 0; ii < nx; ++ii) {
                for (int jj = ny - 1; jj >= 0; --jj)
                    if (UF(s, &C[k], HN, ii, jj) > thr) { yd = jj + 1; break; }
                if (yd > ny - tb) { if ((rc = add_tile(s, n))) return rc; break; }
            }
        }
    }
    /* South */
    size = s->oned ? 0 : s->nactive;
    for (int tt = 0; tt < size; ++tt) {
        int k = s->active[tt], n = s->meta[k].nb[3];
        if (n > -1 && !(in_bounds(s, n) && s->meta[n].on)) {
            long yd = ny;
            for (int ii = 0; ii < nx; ++ii) {
                for (int jj = 0; jj < ny; ++jj)
                    if (UF(s, &C[k], HN, ii, jj) > thr) { yd = jj + 1; break; }
                if (yd <= tb) { if ((rc = add_tile(s, n))) return rc; break; }
            }
        }
    }
    /* East */
    size = s->nactive;
    for (int tt = 0; tt < size; ++tt) {
        int k = s->active[tt], n = s->meta[k].nb[1];
        if (n > -1 && !(in_bounds(s, n) && s->meta[n].on)) {
            long xd = 0;
            for (int jj = 0; jj < ny; ++jj) {
                for (int ii = nx - 1; ii >= 0; --ii)
                    if (UF(s, &C[k], HN, ii, jj) > thr) { xd = ii + 1; break; }
                if (xd > nx - tb) { if ((rc = add_tile(s, n))) return rc; break; }
            }
        }
    }
    /* West */
    size = s->nactive;
    for (int tt = 0; tt < size; ++tt) {
        int k = s->active[tt], n = s->meta[k].nb[0];
        if (n > -1 && !(in_bounds(s, n) && s->meta[n].on)) {
            long xd = nx;
            for (int jj = 0; jj < ny; ++jj) {
                for (int ii = 0; ii < nx; ++ii)
                    if (UF(s, &C[k], HN, ii, jj) > thr) { xd = ii + 1; break; }
                if (xd <= tb) { if ((rc = add_tile(s, n))) return rc; break; }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Time stepping (src/TimeStepper.f90)                                */
/* ------------------------------------------------------------------ */

/* InitialiseTimeSteppingArrays (src/TimeStepper.f90:785-808). */
static void init_stepping_arrays(ko_sim* s) {
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        for (int c = C_I0; c <= C_I3; ++c) tdata_copy(s, &s->cont[c][k], &s->cont[C_TC][k]);
    }
    for (int q = 0; q < s->nghost; ++q) {
        int k = s->ghosts[q];
        for (int c = C_I0; c <= C_I3; ++c) tdata_copy(s, &s->cont[c][k], &s->cont[C_TC][k]);
    }
}

/* CopyMutableTopographicData (src/TimeStepper.f90:838-875). */
static void copy_mutable_topo(ko_sim* s, tdata* from, tdata* to) {
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        tdata *F = &from[k], *T = &to[k];
        int NC = s->NC, NFX = s->NFX, NFY = s->NFY;
        memcpy(&T->u[BT * NC], &F->u[BT * NC], 8 * NC);
        memcpy(&T->u[DBDX * NC], &F->u[DBDX * NC], 8 * NC);
        memcpy(T->bt, F->bt, 8 * s->NV);
        memcpy(&T->upx[BT * NFX], &F->upx[BT * NFX], 8 * NFX);
        memcpy(&T->umx[BT * NFX], &F->umx[BT * NFX], 8 * NFX);
        memcpy(&T->upx[DBDX * NFX], &F->upx[DBDX * NFX], 8 * NFX);
        memcpy(&T->umx[DBDX * NFX], &F->umx[DBDX * NFX], 8 * NFX);
        memcpy(&T->u[DBDY * NC], &F->u[DBDY * NC], 8 * NC);
        memcpy(&T->upx[DBDY * NFX], &F->upx[DBDY * NFX], 8 * NFX);
        memcpy(&T->umx[DBDY * NFX], &F->umx[DBDY * NFX], 8 * NFX);
        memcpy(&T->upy[BT * NFY], &F->upy[BT * NFY], 8 * NFY);
        memcpy(&T->umy[BT * NFY], &F->umy[BT * NFY], 8 * NFY);
        memcpy(&T->upy[DBDX * NFY], &F->upy[DBDX * NFY], 8 * NFY);
        memcpy(&T->umy[DBDX * NFY], &F->umy[DBDX * NFY], 8 * NFY);
        memcpy(&T->upy[DBDY * NFY], &F->upy[DBDY * NFY], 8 * NFY);
        memcpy(&T->umy[DBDY * NFY], &F->umy[DBDY * NFY], 8 * NFY);
    }
}

/* CopySolutionData (src/TimeStepper.f90:811-835). */
static void copy_solution(ko_sim* s, tdata* from, tdata* to) {
    copy_mutable_topo(s, from, to);
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        memcpy(to[k].u, from[k].u, 8 * NF * s->NC);
        memcpy(to[k].dte, from[k].dte, 8 * 4 * s->NC);
        memcpy(to[k].dti, from[k].dti, 8 * 4 * s->NC);
        if (s->cfg.morphodynamics) memcpy(to[k].dtbt, from[k].dtbt, 8 * s->NV);
    }
}

/* UpdateMaximum* (src/TimeStepper.f90:1155-1303) on tileContainer. */
static void update_maxima(ko_sim* s, int k, double t) {
    const kst_config* c = &s->cfg;
    tdata* T = &s->cont[C_TC][k];
    tmeta* m = &s->meta[k];
    int NC = s->NC;
    for (int idx = 0; idx < NC; ++idx) {
        double hn = T->u[HN * NC + idx];
        if (hn > c->height_threshold) {
            if (m->tfirst[idx] == -1) m->tfirst[idx] = t;
            if (hn > m->hnmax[idx]) { m->hnmax[idx] = hn; m->hnmax[NC + idx] = t; }
        }
    }
    for (int idx = 0; idx < NC; ++idx) {
        double spd = sqrt(speed2(c, T->u[UU * NC + idx], T->u[VV * NC + idx], T->u[DBDX * NC + idx], T->u[DBDY * NC + idx]));
        if (spd > m->umax[idx] && T->u[HN * NC + idx] > c->height_threshold) {
            m->umax[idx] = spd; m->umax[NC + idx] = t;
        }
    }
    for (int idx = 0; idx < NC; ++idx) {
        double bt = T->u[BT * NC + idx];
        if (bt < 0 && -bt > m->emax[idx]) { m->emax[idx] = -bt; m->emax[NC + idx] = t; }
        if (bt > 0 && bt > m->dmax[idx]) { m->dmax[idx] = bt; m->dmax[NC + idx] = t; }
    }
    for (int idx = 0; idx < NC; ++idx) {
        double psi = T->u[PSI * NC + idx];
        if (T->u[HN * NC + idx] > c->height_threshold && psi > m->psimax[idx]) {
            m->psimax[idx] = psi; m->psimax[NC + idx] = t;
        }
    }
}

/* ApplySpongeLayer (src/TimeStepper.f90:879-920), hydraulic operator: the
 * explicit RHS of an active tile that lacks a neighbour on some side (tile id
 * 0 in the reference's 1-based ids, -1 here: the first tile's west, the
 * last of the first row's south) is damped by SpongeStrength times the
 * stage state.  Tiles on the domain edge are never activated with sponge
 * boundaries (AddTile), so no active tile qualifies; the check is kept. */
static void apply_sponge(ko_sim* s, int k, tdata* I) {
    if (s->cfg.bc != KST_BC_SPONGE) return;
    const tmeta* m = &s->meta[k];
    if (!((!s->oned && (m->nb[2] == -1 || m->nb[3] == -1)) || m->nb[1] == -1 || m->nb[0] == -1)) return;
    const int NC = s->NC;
    for (int v = 0; v < 4; ++v)
        for (int i = 0; i < NC; ++i)
            I[k].dte[v * NC + i] = I[k].dte[v * NC + i] - s->cfg.sponge_strength * I[k].u[v * NC + i];
}

/* HydraulicTimeStepper (src/TimeStepper.f90:333-527). */
static double hydraulic_stepper(ko_sim* s, double* thisdt, int* refine, int64_t* nrhs) {
    const kst_config* c = &s->cfg;
    tdata *I0 = s->cont[C_I0], *I1 = s->cont[C_I1], *I2 = s->cont[C_I2], *I3 = s->cont[C_I3];
    int NC = s->NC;
    double dt = *thisdt;
    double nextT = s->t + dt;
    s->dt = dt;
    *refine = 0;
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        apply_sponge(s, k, I0);
        for (int v = 0; v < 4; ++v) {
            double *u1 = &I1[k].u[v * NC], *u0 = &I0[k].u[v * NC];
            double *e = &I0[k].dte[v * NC], *im = &I0[k].dti[v * NC];
            for (int i = 0; i < NC; ++i) {
                if (v == RHU || v == RHV) u1[i] = (u0[i] + dt * e[i]) / (1.0 - dt * im[i]);
                else u1[i] = u0[i] + dt * e[i];
            }
        }
    }
    double dt1 = hydraulic_rhs(s, I1, nextT, 2);
    (*nrhs)++;
    if (dt1 < dt) { *thisdt = 0.9 * dt1; *refine = 1; return nextT; }
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        apply_sponge(s, k, I1);
        for (int v = 0; v < 4; ++v) {
            double *u2 = &I2[k].u[v * NC], *u0 = &I0[k].u[v * NC], *u1 = &I1[k].u[v * NC];
            double *e = &I1[k].dte[v * NC], *im = &I1[k].dti[v * NC];
            for (int i = 0; i < NC; ++i) {
                if (v == RHU || v == RHV) {
                    u2[i] = 0.75 * u0[i] + 0.25 * (u1[i] + dt * e[i]) / (1.0 - dt * im[i]);
                } else if (v == W) {
                    double hpo = -I0[k].u[BT * NC + i];
                    hpo = hpo + (I0[k].u[W * NC + i] - I0[k].u[B0 * NC + i]);
                    double hpn = -I1[k].u[BT * NC + i];
                    hpn = hpn + (I1[k].u[W * NC + i] - I1[k].u[B0 * NC + i]);
                    double wu = I0[k].u[BT * NC + i];
                    wu = wu + 0.25 * hpn;
                    wu = wu + 0.75 * hpo;
                    wu = wu + 0.25 * dt * e[i];
                    wu = wu + I0[k].u[B0 * NC + i];
                    u2[i] = wu;
                } else {
                    u2[i] = 0.75 * u0[i] + 0.25 * (u1[i] + dt * e[i]);
                }
            }
        }
    }
    double dt2 = hydraulic_rhs(s, I2, s->t + 0.5 * dt, 3);
    (*nrhs)++;
    if (dt2 < dt) { *thisdt = 0.9 * dt2; *refine = 1; return nextT; }
    const double third = 1.0 / 3.0, twothird = 2.0 / 3.0;
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        apply_sponge(s, k, I2);
        for (int v = 0; v < 4; ++v) {
            double *u3 = &I3[k].u[v * NC], *u0 = &I0[k].u[v * NC], *u2 = &I2[k].u[v * NC];
            double *e = &I2[k].dte[v * NC], *im = &I2[k].dti[v * NC];
            for (int i = 0; i < NC; ++i) {
                if (v == RHU || v == RHV) {
                    u3[i] = third * u0[i] + twothird * (u2[i] + dt * e[i]) / (1.0 - dt * im[i]);
                } else if (v == W) {
                    double hpo = -I0[k].u[BT * NC + i];
                    hpo = hpo + (I0[k].u[W * NC + i] - I0[k].u[B0 * NC + i]);
                    double hpn = -I2[k].u[BT * NC + i];
                    hpn = hpn + (I2[k].u[W * NC + i] - I2[k].u[B0 * NC + i]);
                    double wu = I0[k].u[BT * NC + i];
                    wu = wu + third * hpo;
                    wu = wu + twothird * hpn;
                    wu = wu + twothird * dt * e[i];
                    wu = wu + I0[k].u[B0 * NC + i];
                    u3[i] = wu;
                } else {
                    u3[i] = third * u0[i] + twothird * (u2[i] + dt * e[i]);
                }
            }
        }
    }
    hydraulic_rhs(s, I3, nextT, 4);
    (*nrhs)++;
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        for (int v = RHU; v <= RHV; ++v) {
            double* u3 = &I3[k].u[v * NC];
            double *e = &I3[k].dte[v * NC], *im = &I3[k].dti[v * NC];
            for (int i = 0; i < NC; ++i)
                u3[i] = (u3[i] - dt * dt * e[i] * im[i]) / (1.0 + dt * dt * im[i] * im[i]);
        }
        desingularise(s, s->cont[C_TC], k, 1);
        if (c->track_maxima) update_maxima(s, k, nextT);
    }
    return nextT;
}

/* NextFluxSeriesTime (src/TimeStepper.f90:281-305). */
static double next_flux_series_time(ko_sim* s, double t) {
    double nextT = HUGE_WP, tdiff = HUGE_WP;
    for (int i = 0; i < s->cfg.n_sources; ++i)
        for (int j = 0; j < s->src[i].n_series; ++j) {
            double tmp = s->src[i].time[j] - t;
            if (tmp > 0.0 && tmp < tdiff) { tdiff = tmp; nextT = s->src[i].time[j]; }
        }
    return nextT;
}

/* ------------------------------------------------------------------ */
/* Morphodynamics (src/MorphodynamicRHS.f90, src/TimeStepper.f90:532-781,
 * src/Redistribute.f90)                                              */
/* ------------------------------------------------------------------ */

static double kahan4(double a, double b, double c, double d) {
    double x[4] = {a, b, c, d};
    return kahan_sum(x, 4);
}
static double kahan3(double a, double b, double c) {
    double x[3] = {a, b, c};
    return kahan_sum(x, 3);
}

#define EMD(s, T, i, j) (T)->emd[CI(s, i, j)]
#define DTBT(s, T, i, j) (T)->dtbt[VI(s, i, j)]

/* BtSourceTerm (src/MorphodynamicRHS.f90:154-305), 2D, including the
 * three-term KahanSums of the west edge (:210) and the NE-corner dbdx (:266). */
/* BtSourceTerm on a 1D grid (src/MorphodynamicRHS.f90:272-302): two-cell
 * averages along the row; at a tile edge without an active neighbour the
 * cell's own half. */
static void bt_source_1d(ko_sim* s, tdata* C) {
    const kst_config* c = &s->cfg;
    int nx = s->nx;
    double psib = 1.0 - c->bed_porosity;
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        tdata* T = &C[k];
        tmeta* m = &s->meta[k];
        for (int i = 1; i < nx; ++i) {
            double bx = 0.5 * (UF(s, T, DBDX, i - 1, 0) + UF(s, T, DBDX, i, 0));
            double gam = gamma_of(c, bx, 0.0);
            DTBT(s, T, i, 0) = -0.5 * gam * (EMD(s, T, i - 1, 0) + EMD(s, T, i, 0)) / psib;
        }
        int tw = m->nb[0];
        if (tile_on(s, tw)) {
            double bx = 0.5 * (UF(s, &C[tw], DBDX, nx - 1, 0) + UF(s, T, DBDX, 0, 0));
            double gam = gamma_of(c, bx, 0.0);
            DTBT(s, T, 0, 0) = -0.5 * gam * (EMD(s, &C[tw], nx - 1, 0) + EMD(s, T, 0, 0)) / psib;
        } else {
            double bx = 0.5 * UF(s, T, DBDX, 0, 0);
            double gam = gamma_of(c, bx, 0.0);
            DTBT(s, T, 0, 0) = -0.5 * gam * EMD(s, T, 0, 0) / psib;
        }
        int te = m->nb[1];
        if (tile_on(s, te)) {
            double bx = 0.5 * (UF(s, T, DBDX, nx - 1, 0) + UF(s, &C[te], DBDX, 0, 0));
            double gam = gamma_of(c, bx, 0.0);
            DTBT(s, T, nx, 0) = -0.5 * gam * (EMD(s, T, nx - 1, 0) + EMD(s, &C[te], 0, 0)) / psib;
        } else {
            double bx = 0.5 * UF(s, T, DBDX, nx - 1, 0);
            double gam = gamma_of(c, bx, 0.0);
            DTBT(s, T, nx, 0) = -0.5 * gam * EMD(s, T, nx - 1, 0) / psib;
        }
    }
}

static void bt_source(ko_sim* s, tdata* C) {
    const kst_config* c = &s->cfg;
    int nx = s->nx, ny = s->ny;
    double psib = 1.0 - c->bed_porosity;
    if (s->oned) {
        bt_source_1d(s, C);
        return;
    }
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        tdata* T = &C[k];
        tmeta* m = &s->meta[k];
        for (int i = 1; i < nx; ++i)
            for (int j = 1; j < ny; ++j) {
                double bx = 0.25 * kahan4(UF(s, T, DBDX, i - 1, j - 1), UF(s, T, DBDX, i - 1, j), UF(s, T, DBDX, i, j - 1), UF(s, T, DBDX, i, j));
                double by = 0.25 * kahan4(UF(s, T, DBDY, i - 1, j - 1), UF(s, T, DBDY, i - 1, j), UF(s, T, DBDY, i, j - 1), UF(s, T, DBDY, i, j));
                double gam = gamma_of(c, bx, by);
                DTBT(s, T, i, j) = -0.25 * gam / psib *
                    kahan4(EMD(s, T, i - 1, j - 1), EMD(s, T, i - 1, j), EMD(s, T, i, j - 1), EMD(s, T, i, j));
            }
        tdata* W = &C[m->nb[0]];
        tdata* E = &C[m->nb[1]];
        tdata* N = &C[m->nb[2]];
        tdata* S = &C[m->nb[3]];
        tdata* SW = &C[s->meta[m->nb[0]].nb[3]];
        tdata* SE = &C[s->meta[m->nb[1]].nb[3]];
        tdata* NW = &C[s->meta[m->nb[0]].nb[2]];
        tdata* NE = &C[s->meta[m->nb[1]].nb[2]];
        for (int j = 1; j < ny; ++j) {
            double bx = 0.25 * kahan4(UF(s, W, DBDX, nx - 1, j - 1), UF(s, W, DBDX, nx - 1, j), UF(s, T, DBDX, 0, j - 1), UF(s, T, DBDX, 0, j));
            double by = 0.25 * kahan4(UF(s, W, DBDY, nx - 1, j - 1), UF(s, W, DBDY, nx - 1, j), UF(s, T, DBDY, 0, j - 1), UF(s, T, DBDY, 0, j));
            double gam = gamma_of(c, bx, by);
            DTBT(s, T, 0, j) = -0.25 * gam / psib *
                kahan3(EMD(s, W, nx - 1, j - 1), EMD(s, W, nx - 1, j), EMD(s, T, 0, j - 1) + EMD(s, T, 0, j));
            bx = 0.25 * kahan4(UF(s, T, DBDX, nx - 1, j - 1), UF(s, T, DBDX, nx - 1, j), UF(s, E, DBDX, 0, j - 1), UF(s, E, DBDX, 0, j));
            by = 0.25 * kahan4(UF(s, T, DBDY, nx - 1, j - 1), UF(s, T, DBDY, nx - 1, j), UF(s, E, DBDY, 0, j - 1), UF(s, E, DBDY, 0, j));
            gam = gamma_of(c, bx, by);
            DTBT(s, T, nx, j) = -0.25 * gam / psib *
                kahan4(EMD(s, T, nx - 1, j - 1), EMD(s, T, nx - 1, j), EMD(s, E, 0, j - 1), EMD(s, E, 0, j));
        }
        for (int i = 1; i < nx; ++i) {
            double bx = 0.25 * kahan4(UF(s, T, DBDX, i - 1, ny - 1), UF(s, N, DBDX, i - 1, 0), UF(s, T, DBDX, i, ny - 1), UF(s, N, DBDX, i, 0));
            double by = 0.25 * kahan4(UF(s, T, DBDY, i - 1, ny - 1), UF(s, N, DBDY, i - 1, 0), UF(s, T, DBDY, i, ny - 1), UF(s, N, DBDY, i, 0));
            double gam = gamma_of(c, bx, by);
            DTBT(s, T, i, ny) = -0.25 * gam / psib *
                kahan4(EMD(s, T, i - 1, ny - 1), EMD(s, N, i - 1, 0), EMD(s, T, i, ny - 1), EMD(s, N, i, 0));
            bx = 0.25 * kahan4(UF(s, S, DBDX, i - 1, ny - 1), UF(s, T, DBDX, i - 1, 0), UF(s, S, DBDX, i, ny - 1), UF(s, T, DBDX, i, 0));
            by = 0.25 * kahan4(UF(s, S, DBDY, i - 1, ny - 1), UF(s, T, DBDY, i - 1, 0), UF(s, S, DBDY, i, ny - 1), UF(s, T, DBDY, i, 0));
            gam = gamma_of(c, bx, by);
            DTBT(s, T, i, 0) = -0.25 * gam / psib *
                kahan4(EMD(s, S, i - 1, ny - 1), EMD(s, T, i - 1, 0), EMD(s, S, i, ny - 1), EMD(s, T, i, 0));
        }
        /* corners */
        double bx = 0.25 * kahan4(UF(s, SW, DBDX, nx - 1, ny - 1), UF(s, W, DBDX, nx - 1, 0), UF(s, S, DBDX, 0, ny - 1), UF(s, T, DBDX, 0, 0));
        double by = 0.25 * kahan4(UF(s, SW, DBDY, nx - 1, ny - 1), UF(s, W, DBDY, nx - 1, 0), UF(s, S, DBDY, 0, ny - 1), UF(s, T, DBDY, 0, 0));
        double gam = gamma_of(c, bx, by);
        DTBT(s, T, 0, 0) = -0.25 * gam / psib *
            kahan4(EMD(s, SW, nx - 1, ny - 1), EMD(s, W, nx - 1, 0), EMD(s, S, 0, ny - 1), EMD(s, T, 0, 0));
        bx = 0.25 * kahan4(UF(s, S, DBDX, nx - 1, ny - 1), UF(s, T, DBDX, nx - 1, 0), UF(s, SE, DBDX, 0, ny - 1), UF(s, E, DBDX, 0, 0));
        by = 0.25 * kahan4(UF(s, S, DBDY, nx - 1, ny - 1), UF(s, T, DBDY, nx - 1, 0), UF(s, SE, DBDY, 0, ny - 1), UF(s, E, DBDY, 0, 0));
        gam = gamma_of(c, bx, by);
        DTBT(s, T, nx, 0) = -0.25 * gam / psib *
            kahan4(EMD(s, S, nx - 1, ny - 1), EMD(s, T, nx - 1, 0), EMD(s, SE, 0, ny - 1), EMD(s, E, 0, 0));
        bx = 0.25 * kahan4(UF(s, W, DBDX, nx - 1, ny - 1), UF(s, NW, DBDX, nx - 1, 0), UF(s, T, DBDX, 0, ny - 1), UF(s, N, DBDX, 0, 0));
        by = 0.25 * kahan4(UF(s, W, DBDY, nx - 1, ny - 1), UF(s, NW, DBDY, nx - 1, 0), UF(s, T, DBDY, 0, ny - 1), UF(s, N, DBDY, 0, 0));
        gam = gamma_of(c, bx, by);
        DTBT(s, T, 0, ny) = -0.25 * gam / psib *
            kahan4(EMD(s, W, nx - 1, ny - 1), EMD(s, NW, nx - 1, 0), EMD(s, T, 0, ny - 1), EMD(s, N, 0, 0));
        bx = 0.25 * kahan3(UF(s, T, DBDX, nx - 1, ny - 1), UF(s, N, DBDX, nx - 1, 0), UF(s, E, DBDX, 0, ny - 1) + UF(s, NE, DBDX, 0, 0));
        by = 0.25 * kahan4(UF(s, T, DBDY, nx - 1, ny - 1), UF(s, N, DBDY, nx - 1, 0), UF(s, E, DBDY, 0, ny - 1), UF(s, NE, DBDY, 0, 0));
        gam = gamma_of(c, bx, by);
        DTBT(s, T, nx, ny) = -0.25 * gam / psib *
            kahan4(EMD(s, T, nx - 1, ny - 1), EMD(s, N, nx - 1, 0), EMD(s, E, 0, ny - 1), EMD(s, NE, 0, 0));
    }
}

/* CalculateMorphodynamicRHS (src/MorphodynamicRHS.f90:72-149), 2D. */
static void morpho_rhs(ko_sim* s, tdata* C) {
    const kst_config* c = &s->cfg;
    int nx = s->nx, ny = s->ny;
    double eps = c->height_threshold;
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        tdata* T = &C[k];
        tmeta* m = &s->meta[k];
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                double hW = (i == 0) ? UF(s, &C[m->nb[0]], HN, nx - 1, j) : UF(s, T, HN, i - 1, j);
                double hE = (i == nx - 1) ? UF(s, &C[m->nb[1]], HN, 0, j) : UF(s, T, HN, i + 1, j);
                /* the S and N neighbours only on a 2D grid (src/MorphodynamicRHS.f90:113-135) */
                double hS = eps, hN = eps;
                if (!s->oned) {
                    hS = (j == 0) ? UF(s, &C[m->nb[3]], HN, i, ny - 1) : UF(s, T, HN, i, j - 1);
                    hN = (j == ny - 1) ? UF(s, &C[m->nb[2]], HN, i, 0) : UF(s, T, HN, i, j + 1);
                }
                if (UF(s, T, HN, i, j) < eps || hW < eps || hE < eps || hS < eps || hN < eps) {
                    for (int d = 0; d < 4; ++d) T->dte[d * s->NC + CI(s, i, j)] = 0.0;
                    EMD(s, T, i, j) = 0.0;
                } else {
                    cstate cs = {UF(s, T, HN, i, j), UF(s, T, UU, i, j), UF(s, T, VV, i, j), UF(s, T, PSI, i, j),
                                 UF(s, T, BT, i, j), UF(s, T, DBDX, i, j), UF(s, T, DBDY, i, j)};
                    double E, D;
                    erosion_deposition(c, &cs, &E, &D);
                    EMD(s, T, i, j) = E - D;
                }
            }
    }
    bt_source(s, C);
}

/* ComputeCellCentredTopographicData_ij (src/MorphodynamicRHS.f90:363-416). */
static void cell_topo_ij(ko_sim* s, tdata* T, int i, int j) {
    if (s->oned) {
        UF(s, T, BT, i, 0) = 0.5 * (BTV(s, T, i, 0) + BTV(s, T, i + 1, 0));
        double a[4] = {B0V(s, T, i + 1, 0), BTV(s, T, i + 1, 0), -B0V(s, T, i, 0), -BTV(s, T, i, 0)};
        UF(s, T, DBDX, i, 0) = s->dxr * kahan_sum(a, 4);
        return;
    }
    double bt = 0.25 * kahan4(BTV(s, T, i, j), BTV(s, T, i + 1, j), BTV(s, T, i, j + 1), BTV(s, T, i + 1, j + 1));
    UF(s, T, BT, i, j) = bt;
    double a[8];
    a[0] = B0V(s, T, i + 1, j); a[1] = BTV(s, T, i + 1, j);
    a[2] = -B0V(s, T, i, j); a[3] = -BTV(s, T, i, j);
    a[4] = B0V(s, T, i + 1, j + 1); a[5] = BTV(s, T, i + 1, j + 1);
    a[6] = -B0V(s, T, i, j + 1); a[7] = -BTV(s, T, i, j + 1);
    UF(s, T, DBDX, i, j) = 0.5 * s->dxr * kahan_sum(a, 8);
    a[0] = B0V(s, T, i, j + 1); a[1] = BTV(s, T, i, j + 1);
    a[2] = -B0V(s, T, i, j); a[3] = -BTV(s, T, i, j);
    a[4] = B0V(s, T, i + 1, j + 1); a[5] = BTV(s, T, i + 1, j + 1);
    a[6] = -B0V(s, T, i + 1, j); a[7] = -BTV(s, T, i + 1, j);
    UF(s, T, DBDY, i, j) = 0.5 * s->dyr * kahan_sum(a, 8);
}

/* ExcessDeposition (src/Redistribute.f90:158-198). */
static void excess_deposition(ko_sim* s, int k, int i, int j, double* excess, double* dbt, double* psiold) {
    const kst_config* c = &s->cfg;
    tdata* B = &s->cont[C_I0][k];
    tdata* A = &s->cont[C_I3][k];
    double gamold = gam_cell(s, B, i, j);
    double gamnew = gam_cell(s, A, i, j);
    double hnold = compute_hn(UF(s, B, W, i, j), UF(s, B, B0, i, j), UF(s, B, BT, i, j), gamold);
    (void)gamnew;
    *dbt = UF(s, A, BT, i, j) - UF(s, B, BT, i, j);
    *psiold = UF(s, B, PSI, i, j);
    double e = -(UF(s, B, HNPSI, i, j) * gamold / (1.0 - c->bed_porosity) - *dbt);
    *excess = fmax_(e, -(hnold * gamold - *dbt));
}

typedef struct { int k, i, j; double excess; } rnode;

/* AddToRedistList (src/Redistribute.f90:69-101): ascending, stable. */
static void redist_add(rnode** list, int* n, int* cap, int k, int i, int j, double excess) {
    if (*n == *cap) {
        *cap = *cap ? 2 * *cap : 64;
        *list = realloc(*list, *cap * sizeof(rnode));
    }
    int p = *n;
    while (p > 0 && (*list)[p - 1].excess > excess) { (*list)[p] = (*list)[p - 1]; --p; }
    (*list)[p].k = k; (*list)[p].i = i; (*list)[p].j = j; (*list)[p].excess = excess;
    (*n)++;
}

/* RedistributeCell (src/Redistribute.f90:249-475), 2D. */
static int redistribute_cell(ko_sim* s, int k, int i, int j, double corr) {
    const kst_config* c = &s->cfg;
    tdata* I0 = s->cont[C_I0];
    tdata* I3 = s->cont[C_I3];
    int nx = s->nx, ny = s->ny;
    double bd[4] = {0, 0, 0, 0}, sum = 0.0;
    int di[4], dj[4], N = 0;
    int vi[4] = {i, i + 1, i, i + 1}, vj[4] = {j, j, j + 1, j + 1};
    for (int q = 0; q < 4; ++q) {
        bd[N] = BTV(s, &I3[k], vi[q], vj[q]) - BTV(s, &I0[k], vi[q], vj[q]);
        if (bd[N] > 0.0) {
            di[N] = vi[q]; dj[N] = vj[q];
            sum = sum + bd[N];
            N++;
        }
    }
    if (N == 0) return 0;
    double delta = 4.0 * corr / sum;
    double hnold = UF(s, &I0[k], HN, i, j);
    double gamold = gam_cell(s, &I0[k], i, j);
    double db = UF(s, &I3[k], BT, i, j) - UF(s, &I0[k], BT, i, j);
    double tol = EPS_WP * UF(s, &I3[k], W, i, j) * 10.0;
    double adj = 0.0;
    double disc = kahan3(hnold * gamold, -db, corr);
    if (fabs(disc) < tol) {
        adj = kahan3(tol, -hnold * gamold, db);
        adj = adj * (4.0 / sum);
        adj = adj - delta;
        adj = fmax_(adj, 0.0);
    }
    double hgpo = UF(s, &I0[k], HNPSI, i, j) * gamold / (1.0 - c->bed_porosity);
    tol = EPS_WP * 10.0;
    disc = kahan3(hgpo, -db, corr);
    if (fabs(disc) < tol) {
        double ap = kahan3(tol, -hgpo, db);
        ap = ap * (4.0 / sum);
        ap = ap - delta;
        adj = fmax_(adj, ap);
    }
    delta = delta + adj;
    for (int q = 0; q < N; ++q) {
        double nv = BTV(s, &I3[k], di[q], dj[q]) - delta * bd[q];
        BTV(s, &I3[k], di[q], dj[q]) = nv;
        tmeta* m = &s->meta[k];
        if (di[q] == 0 || di[q] == nx || dj[q] == 0 || dj[q] == ny) {
            int nb[3], nn = 0;
            if (di[q] == 0) nb[nn++] = m->nb[0];
            else if (di[q] == nx) nb[nn++] = m->nb[1];
            if (s->oned) { /* no row copies on a 1D grid */ }
            else if (dj[q] == 0) {
                nb[nn++] = m->nb[3];
                if (di[q] == 0) nb[nn++] = s->meta[m->nb[3]].nb[0];
                else if (di[q] == nx) nb[nn++] = s->meta[m->nb[3]].nb[1];
            } else if (dj[q] == ny) {
                nb[nn++] = m->nb[2];
                if (di[q] == 0) nb[nn++] = s->meta[m->nb[2]].nb[0];
                else if (di[q] == nx) nb[nn++] = s->meta[m->nb[2]].nb[1];
            }
            for (int r = 0; r < nn; ++r)
                if (!s->meta[nb[r]].on) s->redist_ghost++;
        }
        if (di[q] == 0) BTV(s, &I3[m->nb[0]], nx, dj[q]) = nv;
        else if (di[q] == nx) BTV(s, &I3[m->nb[1]], 0, dj[q]) = nv;
        if (s->oned) continue;   /* src/Redistribute.f90:369 */
        if (dj[q] == 0) {
            BTV(s, &I3[m->nb[3]], di[q], ny) = nv;
            if (di[q] == 0) BTV(s, &I3[s->meta[m->nb[3]].nb[0]], nx, ny) = nv;
            if (di[q] == nx) BTV(s, &I3[s->meta[m->nb[3]].nb[1]], 0, ny) = nv;
        }
        if (dj[q] == ny) {
            BTV(s, &I3[m->nb[2]], di[q], 0) = nv;
            if (di[q] == 0) BTV(s, &I3[s->meta[m->nb[2]].nb[0]], nx, 0) = nv;
            if (di[q] == nx) BTV(s, &I3[s->meta[m->nb[2]].nb[1]], 0, 0) = nv;
        }
    }
    /* the 3x3 cells around (i, j), walking tiles as the reference does */
    int ct, ci, cj;
    if (i == 0) { ct = s->meta[k].nb[0]; ci = nx - 1; }
    else { ct = k; ci = i - 1; }
    if (s->oned) {
        /* the three cells of the row, with the 1D order of the w update
         * (src/Redistribute.f90:444-470) */
        for (int kk = -1; kk <= 1; ++kk) {
            tdata* A = &I3[ct];
            tdata* Bo = &I0[ct];
            if (!s->meta[ct].on) s->redist_ghost++;
            cell_topo_ij(s, A, ci, 0);
            double dbb = UF(s, A, BT, ci, 0) - UF(s, Bo, BT, ci, 0);
            double go = gam_cell(s, Bo, ci, 0);
            double gn = gam_cell(s, A, ci, 0);
            double ho = UF(s, Bo, HN, ci, 0);
            double w = -dbb / gn / gn;
            w = w + ho * go / gn / gn;
            w = w + UF(s, A, BT, ci, 0);
            w = w + UF(s, A, B0, ci, 0);
            UF(s, A, W, ci, 0) = w;
            double hpo = UF(s, Bo, HNPSI, ci, 0);
            UF(s, A, HNPSI, ci, 0) = hpo * go / gn - (1.0 - c->bed_porosity) * dbb / gn;
            ci = ci + 1;
            if (ci == nx) { ct = s->meta[ct].nb[1]; ci = 0; }
        }
        return 1;
    }
    for (int kk = -1; kk <= 1; ++kk) {
        if (j == 0) { ct = s->meta[ct].nb[3]; cj = ny - 1; }
        else cj = j - 1;
        for (int l = -1; l <= 1; ++l) {
            tdata* A = &I3[ct];
            tdata* Bo = &I0[ct];
            if (!s->meta[ct].on) s->redist_ghost++;
            cell_topo_ij(s, A, ci, cj);
            double dbb = UF(s, A, BT, ci, cj) - UF(s, Bo, BT, ci, cj);
            double go = gam_cell(s, Bo, ci, cj);
            double gn = gam_cell(s, A, ci, cj);
            double ho = UF(s, Bo, HN, ci, cj);
            double w = UF(s, A, BT, ci, cj);
            w = w + (ho * go / gn - dbb / gn) / gn;
            w = w + UF(s, A, B0, ci, cj);
            UF(s, A, W, ci, cj) = w;
            double hpo = UF(s, Bo, HNPSI, ci, cj);
            UF(s, A, HNPSI, ci, cj) = hpo * go / gn - (1.0 - c->bed_porosity) * dbb / gn;
            cj = cj + 1;
            if (cj == ny && l < 1) { ct = s->meta[ct].nb[2]; cj = 0; }
        }
        if (j + 1 >= ny) ct = s->meta[ct].nb[3];
        ci = ci + 1;
        if (ci == nx) { ct = s->meta[ct].nb[1]; ci = 0; }
    }
    return 1;
}

static void probe_copy(ko_sim* s, tdata** dst, tdata* src) {
    if (!*dst) *dst = xcalloc(s->ntiles, sizeof(tdata));
    for (int k = 0; k < s->ntiles; ++k) {
        if (!src[k].alloc) continue;
        if (!(*dst)[k].alloc) tdata_alloc(s, &(*dst)[k]);
        tdata_copy(s, &(*dst)[k], &src[k]);
    }
}

static void probe_capture(ko_sim* s, int after, const rnode* list, int nl) {
    if (!after) {
        s->probe_dt = s->probe_dt_pending;
    probe_copy(s, &s->probe[0], s->cont[C_I0]);
        probe_copy(s, &s->probe[1], s->cont[C_I3]);
        free(s->probe_list);
        s->probe_list = xcalloc(nl, sizeof(rnode));
        memcpy(s->probe_list, list, nl * sizeof(rnode));
        s->probe_n = nl;
        if (!s->probe_ts) s->probe_ts = xcalloc(s->ntiles, 1);
        for (int k = 0; k < s->ntiles; ++k)
            s->probe_ts[k] = s->meta[k].on ? KST_TILE_ACTIVE : (s->meta[k].ghost ? KST_TILE_GHOST : KST_TILE_OFF);
    } else {
        probe_copy(s, &s->probe[2], s->cont[C_I3]);
        s->probe_done = 1;
    }
}

/* RedistributeGrid (src/Redistribute.f90:203-244): the sorted list in order,
 * re-evaluating each cell's excess; 0 if a cell could not be corrected. */
static int redistribute_grid(ko_sim* s, const rnode* list, int nl) {
    for (int q = 0; q < nl; ++q) {
        double ex, dbt, psio;
        excess_deposition(s, list[q].k, list[q].i, list[q].j, &ex, &dbt, &psio);
        if (ex > EPS_WP) {
            if (!redistribute_cell(s, list[q].k, list[q].i, list[q].j, ex)) return 0;
            s->redist_cells++;
        }
    }
    return 1;
}

/* One linear update of w and Hnpsi from the change of bt
 * (src/TimeStepper.f90:587-610). */
static void morpho_mass_update(ko_sim* s, tdata* I0, tdata* Ik, int k) {
    const kst_config* c = &s->cfg;
    cell_topo(s, &Ik[k]);
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j) {
            double go = gam_cell(s, &I0[k], i, j);
            double gn = gam_cell(s, &Ik[k], i, j);
            double ho = compute_hn(UF(s, &I0[k], W, i, j), UF(s, &I0[k], B0, i, j), UF(s, &I0[k], BT, i, j), go);
            double db = UF(s, &Ik[k], BT, i, j) - UF(s, &I0[k], BT, i, j);
            double w = -db / gn / gn;
            w = w + UF(s, &Ik[k], BT, i, j);
            w = w + ho * go / gn / gn;
            w = w + UF(s, &Ik[k], B0, i, j);
            UF(s, &Ik[k], W, i, j) = w;
            double hpo = UF(s, &I0[k], HNPSI, i, j);
            double hp = -(1.0 - c->bed_porosity) * db / gn;
            hp = hp + hpo * go / gn;
            UF(s, &Ik[k], HNPSI, i, j) = hp;
        }
    desingularise(s, Ik, k, 0);
}

/* MorphodynamicTimeStepper (src/TimeStepper.f90:532-781).  Its sponge-layer
 * calls (ApplySpongeLayer with morpho_flag, :572,618,663) damp the bt rate of
 * active tiles lacking a neighbour, which no active tile does (see
 * apply_sponge): they are left out. */
static void morpho_stepper(ko_sim* s, double* thisdt, int* refine) {
    const kst_config* c = &s->cfg;
    tdata *I0 = s->cont[C_I0], *I1 = s->cont[C_I1], *I2 = s->cont[C_I2], *I3 = s->cont[C_I3];
    int NV = s->NV;
    double dt = *thisdt;
    *refine = 0;
    double relmax = 0.0;
    s->probe_dt_pending = dt;
    morpho_rhs(s, I0);
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        for (int v = 0; v < NV; ++v) I1[k].bt[v] = fmax_(-c->ero_depth, I0[k].bt[v] + dt * I0[k].dtbt[v]);
    }
    for (int q = 0; q < s->nactive; ++q) morpho_mass_update(s, I0, I1, s->active[q]);
    morpho_rhs(s, I1);
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        for (int v = 0; v < NV; ++v)
            I2[k].bt[v] = fmax_(-c->ero_depth, 0.75 * I0[k].bt[v] + 0.25 * (I1[k].bt[v] + dt * I1[k].dtbt[v]));
    }
    for (int q = 0; q < s->nactive; ++q) morpho_mass_update(s, I0, I2, s->active[q]);
    morpho_rhs(s, I2);
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        for (int v = 0; v < NV; ++v)
            I3[k].bt[v] = fmax_(-c->ero_depth, (1.0 / 3.0) * I0[k].bt[v] + (2.0 / 3.0) * (I2[k].bt[v] + dt * I2[k].dtbt[v]));
    }
    for (int q = 0; q < s->nactive; ++q) morpho_mass_update(s, I0, I3, s->active[q]);

    rnode* list = NULL;
    int nl = 0, cap = 0;
    rnode* scan = NULL;       /* the same candidates in scan order (probe only) */
    int ns = 0, scap = 0;
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                double go = gam_cell(s, &I0[k], i, j);
                double gn = gam_cell(s, &I3[k], i, j);
                double ho = compute_hn(UF(s, &I0[k], W, i, j), UF(s, &I0[k], B0, i, j), UF(s, &I0[k], BT, i, j), go);
                double hn = compute_hn(UF(s, &I3[k], W, i, j), UF(s, &I3[k], B0, i, j), UF(s, &I3[k], BT, i, j), gn);
                double ex, dbt, psio;
                excess_deposition(s, k, i, j, &ex, &dbt, &psio);
                if (!*refine && ex > EPS_WP && dbt > EPS_WP && psio > -EPS_WP) {
                    if (ho < c->ero_critical_height || fabs(hn - ho) < c->ero_critical_height) {
                        redist_add(&list, &nl, &cap, k, i, j, ex);
                        if (s->probe_on && !s->probe_done) {
                            if (ns == scap) { scap = scap ? 2 * scap : 64; scan = realloc(scan, scap * sizeof(rnode)); }
                            scan[ns].k = k; scan[ns].i = i; scan[ns].j = j; scan[ns].excess = ex; ns++;
                        }
                    } else {
                        *refine = 1;
                        break;
                    }
                    continue;
                }
                if (ho < c->ero_critical_height) continue;
                double rel = fabs(hn - ho) / fabs(ho);
                if (rel > 0.1 && rel > relmax) {
                    relmax = rel;
                    *refine = 1;
                    break;
                }
            }
    }
    int probing = ((nl > 0 || s->probe_on == 2) && !*refine && s->probe_on && !s->probe_done);
    if (probing) probe_capture(s, 0, scan, ns);
    if (nl > 0 && !*refine && !redistribute_grid(s, list, nl)) *refine = 1;
    if (probing) probe_capture(s, 1, NULL, 0);
    free(list);
    free(scan);
    if (*refine) {
        *thisdt = 0.5 * dt;
        return;
    }
    for (int q = 0; q < s->nactive; ++q) face_topo(s, I3, s->active[q]);
}

static int64_t count_cells(ko_sim* s) { return (int64_t)s->nactive * s->NC; }

/* One accepted outer step of IntegrateTo (src/TimeStepper.f90:151-275):
 * H(dt) alone, or the Strang step H(dt) M(2dt) H(dt) with morphodynamics. */
static int one_step(ko_sim* s, double tend, kst_stats* st) {
    int rc = check_near_boundaries(s);
    if (rc) return rc;
    init_stepping_arrays(s);
    double adv = hydraulic_rhs(s, s->cont[C_TC], s->t, 1);
    st->rhs_evals++;
    double tmax = fmin_(tend, next_flux_series_time(s, s->t));
    if (!s->cfg.morphodynamics) adv = fmin_(adv, tmax - s->t);
    else adv = fmin_(adv, 0.5 * (tmax - s->t));
    s->dt = adv;
    double dth = adv;
    s->t0 = s->t;
    tdata *TC = s->cont[C_TC], *I0 = s->cont[C_I0], *I1 = s->cont[C_I1];
    tdata *I2 = s->cont[C_I2], *I3 = s->cont[C_I3];
    for (;;) {
        copy_solution(s, TC, I0);
        copy_mutable_topo(s, TC, I1);
        copy_mutable_topo(s, TC, I2);
        copy_solution(s, TC, I3);
        int refine = 0;
        hydraulic_stepper(s, &dth, &refine, &st->rhs_evals);
        if (refine) { st->rejected++; continue; }
        if (!s->cfg.morphodynamics) break;
        copy_solution(s, I3, I0);
        copy_solution(s, I3, I1);
        copy_solution(s, I3, I2);
        double dtm = 2.0 * dth;
        morpho_stepper(s, &dtm, &refine);
        if (refine) {
            dth = 0.5 * dtm;
            s->dt = dth;
            st->rejected++;
            continue;
        }
        copy_solution(s, I3, I0);
        copy_mutable_topo(s, I3, I1);
        copy_mutable_topo(s, I3, I2);
        s->t = s->t + dth;
        double adv2 = hydraulic_rhs(s, I0, s->t, 1);
        st->rhs_evals++;
        if (adv2 < dth) {
            if ((dth - adv2) / dth < 0.1) dth = 0.9 * dth;
            else dth = adv2;
            s->dt = dth;
            s->t = s->t0;
            st->rejected++;
            continue;
        }
        hydraulic_stepper(s, &dth, &refine, &st->rhs_evals);
        if (!refine) break;
        s->t = s->t0;
        st->rejected++;
    }
    copy_solution(s, I3, TC);
    if (!s->cfg.morphodynamics) s->t = s->t0 + dth;
    else s->t = s->t0 + 2.0 * dth;
    st->steps++;
    st->cell_updates += count_cells(s);
    st->last_dt = dth;
    if (s->steplog && s->steplog_n < s->steplog_cap) {
        s->steplog[2 * s->steplog_n] = s->t;
        s->steplog[2 * s->steplog_n + 1] = dth;
        s->steplog_n++;
    }
    if (st->steps == 1 || dth < st->min_dt) st->min_dt = dth;
    if (dth > st->max_dt) st->max_dt = dth;
    return 0;
}

static void finish_stats(ko_sim* s, kst_stats* st) {
    st->redist_cells = s->redist_cells;
    st->redist_ghost = s->redist_ghost;
    st->active_tiles = s->nactive;
    st->ghost_tiles = s->nghost;
    st->t = s->t;
}

int ko_integrate_to(ko_sim* s, double tend, kst_stats* st) {
    kst_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof *st);
    if (s->status) return s->status;
    s->redist_cells = s->redist_ghost = 0;
    if (tend <= s->t) { finish_stats(s, st); return 0; }
    for (;;) {
        int rc = one_step(s, tend, st);
        if (rc) return rc;
        if (s->t >= tend) break;
    }
    finish_stats(s, st);
    return 0;
}

int ko_step_n(ko_sim* s, int64_t n, kst_stats* st) {
    kst_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof *st);
    if (s->status) return s->status;
    s->redist_cells = s->redist_ghost = 0;
    for (int64_t q = 0; q < n; ++q) {
        int rc = one_step(s, HUGE_WP, st);
        if (rc) return rc;
    }
    finish_stats(s, st);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Setup and I/O                                                      */
/* ------------------------------------------------------------------ */

ko_sim* ko_create(const kst_config* cfg) {
    ko_sim* s = xcalloc(1, sizeof *s);
    s->cfg = *cfg;
    s->nx = cfg->nx_per_tile; s->ny = cfg->ny_per_tile;
    s->ntx = cfg->nx_tiles; s->nty = cfg->ny_tiles;
    s->ntiles = s->ntx * s->nty;
    s->oned = s->nty * s->ny == 1;
    s->NX = s->ntx * s->nx; s->NY = s->nty * s->ny;
    s->NC = s->nx * s->ny; s->NFX = (s->nx + 1) * s->ny; s->NFY = s->nx * (s->ny + 1);
    s->NV = (s->nx + 1) * (s->ny + 1);
    s->dxr = 1.0 / cfg->dx; s->dyr = 1.0 / cfg->dy;
    s->meta = xcalloc(s->ntiles, sizeof(tmeta));
    for (int c = 0; c < NCONT; ++c) s->cont[c] = xcalloc(s->ntiles, sizeof(tdata));
    s->active = xcalloc(s->ntiles + 1, sizeof(int));
    s->ghosts = xcalloc(s->ntiles + 1, sizeof(int));
    for (int k = 0; k < s->ntiles; ++k) set_tile_boundaries(s, k);
    s->gb0 = xcalloc((size_t)(s->NX + 1) * (s->NY + 1), 8);
    s->dt = 1.0e-5;
    int ns = cfg->n_sources;
    s->src = xcalloc(ns, sizeof(kst_flux_source));
    s->src_arrays = xcalloc(3 * ns, sizeof(double*));
    for (int i = 0; i < ns; ++i) {
        s->src[i] = cfg->sources[i];
        int n = cfg->sources[i].n_series;
        for (int a = 0; a < 3; ++a) {
            const double* from = a == 0 ? cfg->sources[i].time : a == 1 ? cfg->sources[i].flux : cfg->sources[i].psi;
            s->src_arrays[3 * i + a] = xcalloc(n, 8);
            memcpy(s->src_arrays[3 * i + a], from, 8 * n);
        }
        s->src[i].time = s->src_arrays[3 * i];
        s->src[i].flux = s->src_arrays[3 * i + 1];
        s->src[i].psi = s->src_arrays[3 * i + 2];
    }
    s->cfg.sources = s->src;
    if (cfg->bc < KST_BC_HALT || cfg->bc > KST_BC_SPONGE) fail(s, KST_ERR_CONFIG, "unknown boundary condition");
    return s;
}

void ko_destroy(ko_sim* s) {
    if (!s) return;
    for (int c = 0; c < NCONT; ++c) {
        for (int k = 0; k < s->ntiles; ++k) tdata_free(&s->cont[c][k]);
        free(s->cont[c]);
    }
    for (int k = 0; k < s->ntiles; ++k) {
        tmeta* m = &s->meta[k];
        free(m->x); free(m->y); free(m->xv); free(m->yv);
        free(m->hnmax); free(m->umax); free(m->emax); free(m->dmax); free(m->psimax); free(m->tfirst);
    }
    for (int p = 0; p < 3; ++p) {
        if (!s->probe[p]) continue;
        for (int k = 0; k < s->ntiles; ++k) tdata_free(&s->probe[p][k]);
        free(s->probe[p]);
    }
    free(s->probe_list);
    free(s->probe_ts);
    for (int i = 0; i < 3 * s->cfg.n_sources; ++i) free(s->src_arrays[i]);
    free(s->src_arrays); free(s->src);
    free(s->meta); free(s->active); free(s->ghosts); free(s->gb0);
    free(s);
}

const char* ko_error(const ko_sim* s) { return s->err; }

int ko_set_topography(ko_sim* s, const double* b0v) {
    memcpy(s->gb0, b0v, 8 * (size_t)(s->NX + 1) * (s->NY + 1));
    return 0;
}

/* Cap painting of LoadSourceConditions (src/SetSources.f90:228-303), 2D. */
int ko_add_caps(ko_sim* s, int ncaps, const double* cx, const double* cy, const double* cr,
                const double* ch, const double* cpsi, const double* cu, const double* cv,
                const int* shape) {
    const kst_config* c = &s->cfg;
    int nx = s->nx, ny = s->ny;
    for (int gi = 0; gi < s->ntx; ++gi)
        for (int gj = 0; gj < s->nty; ++gj) {
            int k = gi + gj * s->ntx;
            for (int ii = 0; ii < nx; ++ii)
                for (int jj = 0; jj < ny; ++jj) {
                    double x = -0.5 * c->x_size + c->dx * (((gi + 1) - 1.0) * nx + ((ii + 1) - 0.5));
                    double y = -0.5 * c->y_size + c->dy * (((gj + 1) - 1.0) * ny + ((jj + 1) - 0.5));
                    for (int q = 0; q < ncaps; ++q) {
                        double capR = cr[q], capHn = ch[q], cappsi = cpsi[q];
                        double rho = c->rhow + (c->rhos - c->rhow) * cappsi;
                        double Hn = capHn;
                        double R2 = (x - cx[q]) * (x - cx[q]) + (y - cy[q]) * (y - cy[q]);
                        if (R2 <= capR * capR) {
                            if (!s->meta[k].on) {
                                int rc = add_tile(s, k);
                                if (rc) return rc;
                            }
                            tdata* T = &s->cont[C_TC][k];
                            double gam = gam_cell(s, T, ii, jj);
                            if (s->meta[k].on) {
                                double hn0 = UF(s, T, HN, ii, jj), rho0 = UF(s, T, RHO, ii, jj);
                                tmeta* m = &s->meta[k];
                                int idx = CI(s, ii, jj);
                                if (shape[q] == 0) {
                                    UF(s, T, W, ii, jj) = UF(s, T, W, ii, jj) + capHn / gam;
                                    UF(s, T, HN, ii, jj) = UF(s, T, HN, ii, jj) + capHn;
                                    m->hnmax[idx] = m->hnmax[idx] + capHn;
                                    UF(s, T, RHU, ii, jj) = UF(s, T, RHU, ii, jj) + rho * capHn * cu[q];
                                    UF(s, T, RHV, ii, jj) = UF(s, T, RHV, ii, jj) + rho * capHn * cv[q];
                                    UF(s, T, HNPSI, ii, jj) = UF(s, T, HNPSI, ii, jj) + cappsi * capHn;
                                } else if (shape[q] == 1) {
                                    UF(s, T, W, ii, jj) = UF(s, T, W, ii, jj) + capHn * (1.0 - R2 / capR / capR) / gam;
                                    UF(s, T, HN, ii, jj) = UF(s, T, HN, ii, jj) + capHn * (1.0 - R2 / capR / capR);
                                    m->hnmax[idx] = m->hnmax[idx] + capHn * (1.0 - R2 / capR / capR);
                                    UF(s, T, RHU, ii, jj) = UF(s, T, RHU, ii, jj) + rho * capHn * cu[q];
                                    UF(s, T, RHV, ii, jj) = UF(s, T, RHV, ii, jj) + rho * capHn * cv[q];
                                    UF(s, T, HNPSI, ii, jj) = UF(s, T, HNPSI, ii, jj) + cappsi * capHn * (1.0 - R2 / capR / capR);
                                } else {
                                    double hp = capHn - UF(s, T, B0, ii, jj);
                                    Hn = hp * gam;
                                    if (Hn > 0.0) {
                                        UF(s, T, W, ii, jj) = UF(s, T, W, ii, jj) + hp;
                                        UF(s, T, HN, ii, jj) = UF(s, T, HN, ii, jj) + Hn;
                                        m->hnmax[idx] = m->hnmax[idx] + Hn;
                                        UF(s, T, RHU, ii, jj) = UF(s, T, RHU, ii, jj) + rho * Hn * cu[q];
                                        UF(s, T, RHV, ii, jj) = UF(s, T, RHV, ii, jj) + rho * Hn * cv[q];
                                        UF(s, T, HNPSI, ii, jj) = UF(s, T, HNPSI, ii, jj) + cappsi * Hn;
                                    }
                                }
                                UF(s, T, RHO, ii, jj) = (rho0 * hn0 + rho * Hn) / (hn0 + Hn);
                            }
                        }
                    }
                }
        }
    return 0;
}

/* Flux-source cells of LoadSourceConditions (src/SetSources.f90:357-375). */
int ko_add_flux_sources(ko_sim* s, int32_t* num_cells_out) {
    const kst_config* c = &s->cfg;
    int nx = s->nx, ny = s->ny, ns = c->n_sources;
    for (int q = 0; q < ns; ++q) s->src[q].num_cells = 0;
    for (int gi = 0; gi < s->ntx; ++gi)
        for (int gj = 0; gj < s->nty; ++gj) {
            int k = gi + gj * s->ntx;
            for (int ii = 0; ii < nx; ++ii)
                for (int jj = 0; jj < ny; ++jj) {
                    double x = -0.5 * c->x_size + c->dx * (((gi + 1) - 1.0) * nx + ((ii + 1) - 0.5));
                    double y = -0.5 * c->y_size + c->dy * (((gj + 1) - 1.0) * ny + ((jj + 1) - 0.5));
                    for (int q = 0; q < ns; ++q) {
                        double R2 = (x - s->src[q].x) * (x - s->src[q].x) + (y - s->src[q].y) * (y - s->src[q].y);
                        if (R2 <= s->src[q].radius * s->src[q].radius) {
                            if (!s->meta[k].on) {
                                int rc = add_tile(s, k);
                                if (rc) return rc;
                            }
                            s->meta[k].src = 1;
                            s->src[q].num_cells++;
                        }
                    }
                }
        }
    if (num_cells_out)
        for (int q = 0; q < ns; ++q) num_cells_out[q] = s->src[q].num_cells;
    return 0;
}

int ko_finish_sources(ko_sim* s) {
    for (int q = 0; q < s->nghost; ++q) reconstruct_w_at_edges(s, s->ghosts[q]);
    return 0;
}

/* State load with the C-ABI's semantics: activate tiles flagged ACTIVE in id
 * order (AddTile's bookkeeping), mark the flagged ghosts, load the fields,
 * recompute the derived variables and rebuild the ghost w reconstructions. */
int ko_set_state(ko_sim* s, const double* w, const double* rhu, const double* rhv,
                 const double* hnpsi, const double* btv, const uint8_t* ts, double t,
                 const double* hnmax0) {
    for (int k = 0; k < s->ntiles; ++k)
        if (ts[k] == KST_TILE_ACTIVE && !s->meta[k].on) {
            int rc = add_tile(s, k);
            if (rc) return rc;
        }
    int nx = s->nx, ny = s->ny;
    /* containsSource: tiles holding a cell within a flux source's radius
     * (src/SetSources.f90:357-375); NumCellsInSrc comes with the config */
    const kst_config* c = &s->cfg;
    for (int k = 0; k < s->ntiles; ++k) {
        int gi, gj;
        tile_coords(s, k, &gi, &gj);
        for (int q = 0; q < c->n_sources; ++q)
            for (int jj = 0; jj < ny; ++jj)
                for (int ii = 0; ii < nx; ++ii) {
                    double x = -0.5 * c->x_size + c->dx * (((gi + 1) - 1.0) * nx + ((ii + 1) - 0.5));
                    double y = -0.5 * c->y_size + c->dy * (((gj + 1) - 1.0) * ny + ((jj + 1) - 0.5));
                    /* along x only on a 1D grid (src/SetSources.f90:219-231) */
                    double R2 = (x - s->src[q].x) * (x - s->src[q].x);
                    if (!s->oned) R2 = R2 + (y - s->src[q].y) * (y - s->src[q].y);
                    if (R2 <= s->src[q].radius * s->src[q].radius) s->meta[k].src = 1;
                }
    }
    if (btv) {
        /* every active tile takes its vertex bt from the global array (ghost
         * tiles keep bt = 0, as AllocateTopographicData leaves them); the
         * cell and face topography follow */
        for (int k = 0; k < s->ntiles; ++k) {
            tdata* T = &s->cont[C_TC][k];
            if (!T->alloc || !s->meta[k].on) continue;
            int gi, gj;
            tile_coords(s, k, &gi, &gj);
            for (int j = 0; j <= ny; ++j)
                for (int i = 0; i <= nx; ++i)
                    BTV(s, T, i, j) = btv[(size_t)(gj * ny + j) * (s->NX + 1) + gi * nx + i];
            cell_topo(s, T);
        }
        for (int q = 0; q < s->nactive; ++q) face_topo(s, s->cont[C_TC], s->active[q]);
    }
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q], gi, gj;
        tile_coords(s, k, &gi, &gj);
        tdata* T = &s->cont[C_TC][k];
        for (int jj = 0; jj < ny; ++jj)
            for (int ii = 0; ii < nx; ++ii) {
                size_t g = (size_t)(gj * ny + jj) * s->NX + gi * nx + ii;
                UF(s, T, W, ii, jj) = w[g];
                UF(s, T, RHU, ii, jj) = rhu[g];
                UF(s, T, RHV, ii, jj) = rhv[g];
                UF(s, T, HNPSI, ii, jj) = hnpsi[g];
                if (hnmax0) s->meta[k].hnmax[CI(s, ii, jj)] = hnmax0[g];
            }
        desingularise(s, s->cont[C_TC], k, 1);
    }
    s->t = t;
    return ko_finish_sources(s);
}

/* The cell bt, db/dx, db/dy of LoadInitialCondition_txt (src/Restart.f90:216-236)
 * for the active tiles' cells where cells[] is set: ActivateTile(restart)
 * keeps them (src/UpdateTiles.f90:342-384); the derived variables follow
 * (on a 1D grid db/dy is not read). */
int ko_set_cell_topography(ko_sim* s, const double* btc, const double* dbdx, const double* dbdy,
                           const uint8_t* cells) {
    int nx = s->nx, ny = s->ny;
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q], gi, gj;
        tile_coords(s, k, &gi, &gj);
        tdata* T = &s->cont[C_TC][k];
        for (int jj = 0; jj < ny; ++jj)
            for (int ii = 0; ii < nx; ++ii) {
                size_t g = (size_t)(gj * ny + jj) * s->NX + gi * nx + ii;
                if (!cells[g]) continue;
                UF(s, T, BT, ii, jj) = btc[g];
                UF(s, T, DBDX, ii, jj) = dbdx[g];
                if (!s->oned) UF(s, T, DBDY, ii, jj) = dbdy[g];
            }
        desingularise(s, s->cont[C_TC], k, 1);
    }
    return 0;
}

static void scatter_field(ko_sim* s, int k, const double* src, double* dst) {
    int gi, gj, nx = s->nx, ny = s->ny;
    tile_coords(s, k, &gi, &gj);
    for (int jj = 0; jj < ny; ++jj)
        for (int ii = 0; ii < nx; ++ii) dst[(size_t)(gj * ny + jj) * s->NX + gi * nx + ii] = src[CI(s, ii, jj)];
}

int ko_get_state(ko_sim* s, double* w, double* rhu, double* rhv, double* hnpsi, double* hn,
                 double* btv, uint8_t* ts, double* t) {
    size_t N = (size_t)s->NX * s->NY;
    double* outs[5] = {w, rhu, rhv, hnpsi, hn};
    int fld[5] = {W, RHU, RHV, HNPSI, HN};
    for (int a = 0; a < 5; ++a)
        if (outs[a]) memset(outs[a], 0, 8 * N);
    for (int k = 0; k < s->ntiles; ++k) {
        tdata* T = &s->cont[C_TC][k];
        if (!T->alloc) continue;
        for (int a = 0; a < 5; ++a)
            if (outs[a]) scatter_field(s, k, &T->u[fld[a] * s->NC], outs[a]);
    }
    if (btv) {
        /* vertex (I, J) from the tile holding cell (min(I, NX-1), min(J, NY-1)) */
        memset(btv, 0, 8 * (size_t)(s->NX + 1) * (s->NY + 1));
        for (int J = 0; J <= s->NY; ++J)
            for (int I = 0; I <= s->NX; ++I) {
                int ci = I < s->NX ? I : s->NX - 1, cj = J < s->NY ? J : s->NY - 1;
                int k = (cj / s->ny) * s->ntx + ci / s->nx;
                tdata* T = &s->cont[C_TC][k];
                if (!T->alloc) continue;
                btv[(size_t)J * (s->NX + 1) + I] = BTV(s, T, I - (ci / s->nx) * s->nx, J - (cj / s->ny) * s->ny);
            }
    }
    if (ts)
        for (int k = 0; k < s->ntiles; ++k)
            ts[k] = s->meta[k].on ? KST_TILE_ACTIVE : (s->meta[k].ghost ? KST_TILE_GHOST : KST_TILE_OFF);
    if (t) *t = s->t;
    return 0;
}

int ko_get_maxima(ko_sim* s, double* const mx[11]) {
    size_t N = (size_t)s->NX * s->NY;
    for (int a = 0; a < 11; ++a) memset(mx[a], 0, 8 * N);
    for (int a = 0; a < (int)N; ++a) mx[10][a] = -1.0;
    for (int k = 0; k < s->ntiles; ++k) {
        tmeta* m = &s->meta[k];
        if (!m->hnmax) continue;
        double* srcs[5] = {m->hnmax, m->umax, m->emax, m->dmax, m->psimax};
        for (int a = 0; a < 5; ++a) {
            scatter_field(s, k, srcs[a], mx[2 * a]);
            scatter_field(s, k, srcs[a] + s->NC, mx[2 * a + 1]);
        }
        scatter_field(s, k, m->tfirst, mx[10]);
    }
    return 0;
}

/* The inverse of ko_get_maxima for the allocated tiles (maxima painted by
 * the initial conditions, restarts). */
int ko_set_maxima(ko_sim* s, const double* const mx[11]) {
    int gi, gj, nx = s->nx, ny = s->ny;
    for (int k = 0; k < s->ntiles; ++k) {
        tmeta* m = &s->meta[k];
        if (!m->hnmax) continue;
        tile_coords(s, k, &gi, &gj);
        double* dsts[5] = {m->hnmax, m->umax, m->emax, m->dmax, m->psimax};
        for (int jj = 0; jj < ny; ++jj)
            for (int ii = 0; ii < nx; ++ii) {
                size_t g = (size_t)(gj * ny + jj) * s->NX + gi * nx + ii;
                for (int a = 0; a < 5; ++a) {
                    dsts[a][CI(s, ii, jj)] = mx[2 * a][g];
                    dsts[a][s->NC + CI(s, ii, jj)] = mx[2 * a + 1][g];
                }
                m->tfirst[CI(s, ii, jj)] = mx[10][g];
            }
    }
    return 0;
}

int ko_get_fields13(ko_sim* s, double* out) {
    size_t N = (size_t)s->NX * s->NY;
    memset(out, 0, 8 * N * NF);
    for (int k = 0; k < s->ntiles; ++k) {
        tdata* T = &s->cont[C_TC][k];
        if (!T->alloc) continue;
        for (int d = 0; d < NF; ++d) scatter_field(s, k, &T->u[d * s->NC], out + d * N);
    }
    return 0;
}

/* CalculateVolume (src/Output.f90:617-735), 2D: over the active tiles in
 * list order and their cells row by row (j outer, i inner), Kahan sums
 * (src/utilities.f90:439-448) of the flow volume (w - b0 - bt) gam^2, the
 * mass rho Hn gam, the bed change bt and the solids Hnpsi gam, read from
 * tileContainer's stored u (Hn and rho as ComputeDesingularisedVariables
 * left them); each times dx dy, then the Volume.txt columns 2-7. */
int ko_volume(ko_sim* s, double out[6]) {
    const kst_config* c = &s->cfg;
    double vol = 0.0, mass = 0.0, bed = 0.0, solids = 0.0;
    double cv = 0.0, cm = 0.0, cb = 0.0, cs = 0.0;
    for (int q = 0; q < s->nactive; ++q) {
        tdata* T = &s->cont[C_TC][s->active[q]];
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                double x[4];
                const double gam = gam_cell(s, T, i, j);
                const double hn = UF(s, T, HN, i, j), rho = UF(s, T, RHO, i, j);
                x[0] = (UF(s, T, W, i, j) - UF(s, T, B0, i, j) - UF(s, T, BT, i, j)) * gam * gam;
                x[1] = rho * hn * gam;
                x[2] = UF(s, T, BT, i, j);
                x[3] = UF(s, T, HNPSI, i, j) * gam;
                double* acc[4] = {&vol, &mass, &bed, &solids};
                double* cmp[4] = {&cv, &cm, &cb, &cs};
                for (int m = 0; m < 4; ++m) {
                    double y = x[m] - *cmp[m];
                    double t = *acc[m] + y;
                    *cmp[m] = (t - *acc[m]) - y;
                    *acc[m] = t;
                }
            }
    }
    vol = vol * c->dx * c->dy;
    mass = mass * c->dx * c->dy;
    bed = bed * c->dx * c->dy;
    solids = solids * c->dx * c->dy;
    const double rhob = c->rhow * c->bed_porosity + c->rhos * (1.0 - c->bed_porosity);
    out[0] = vol;
    out[1] = bed;
    out[2] = mass;
    out[3] = bed * rhob;
    out[4] = solids * c->rhos;
    out[5] = bed * c->rhos * (1.0 - c->bed_porosity);
    return 0;
}

/* The four per-cell terms of ko_volume in summation order (4 x NC x nactive). */
int ko_volume_terms(ko_sim* s, double* terms, int64_t cap) {
    int64_t n = 0;
    for (int q = 0; q < s->nactive; ++q) {
        tdata* T = &s->cont[C_TC][s->active[q]];
        for (int j = 0; j < s->ny; ++j)
            for (int i = 0; i < s->nx; ++i) {
                if (n >= cap) return -1;
                const double gam = gam_cell(s, T, i, j);
                terms[4 * n + 0] = (UF(s, T, W, i, j) - UF(s, T, B0, i, j) - UF(s, T, BT, i, j)) * gam * gam;
                terms[4 * n + 1] = UF(s, T, RHO, i, j) * UF(s, T, HN, i, j) * gam;
                terms[4 * n + 2] = UF(s, T, BT, i, j);
                terms[4 * n + 3] = UF(s, T, HNPSI, i, j) * gam;
                ++n;
            }
    }
    return (int)n;
}

int ko_get_ghost_slopes(ko_sim* s, double* sx, double* sy) {
    size_t N = (size_t)s->NX * s->NY;
    memset(sx, 0, 8 * N);
    memset(sy, 0, 8 * N);
    for (int q = 0; q < s->nghost; ++q) {
        int k = s->ghosts[q];
        tdata* T = &s->cont[C_TC][k];
        scatter_field(s, k, &T->ulx[W * s->NC], sx);
        scatter_field(s, k, &T->uly[W * s->NC], sy);
    }
    return 0;
}

int ko_rhs(ko_sim* s, int substep, double t, double* ddte, double* ddti, double* advised) {
    double keep = s->dt;
    double adv = hydraulic_rhs(s, s->cont[C_TC], t, substep);
    s->dt = keep;
    size_t N = (size_t)s->NX * s->NY;
    memset(ddte, 0, 8 * 4 * N);
    memset(ddti, 0, 8 * 4 * N);
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        for (int v = 0; v < 4; ++v) {
            scatter_field(s, k, &s->cont[C_TC][k].dte[v * s->NC], ddte + v * N);
            scatter_field(s, k, &s->cont[C_TC][k].dti[v * s->NC], ddti + v * N);
        }
    }
    if (advised) *advised = adv;
    return 0;
}

int ko_tile_lists(ko_sim* s, int32_t* active, int32_t* nactive, int32_t* ghost, int32_t* nghost) {
    if (active) for (int q = 0; q < s->nactive; ++q) active[q] = s->active[q];
    if (ghost) for (int q = 0; q < s->nghost; ++q) ghost[q] = s->ghosts[q];
    if (nactive) *nactive = s->nactive;
    if (nghost) *nghost = s->nghost;
    return 0;
}

/* ---- probes for the tests (no counterpart in the reference) ---- */

static void gather13(ko_sim* s, tdata* C, double* f13, double* btt) {
    size_t N = (size_t)s->NX * s->NY;
    if (f13) memset(f13, 0, 8 * N * NF);
    if (btt) memset(btt, 0, 8 * (size_t)s->ntiles * s->NV);
    for (int k = 0; k < s->ntiles; ++k) {
        tdata* T = &C[k];
        if (!T->alloc) continue;
        if (f13)
            for (int d = 0; d < NF; ++d) scatter_field(s, k, &T->u[d * s->NC], f13 + d * N);
        if (btt) memcpy(btt + (size_t)k * s->NV, T->bt, 8 * s->NV);
    }
}

/* Per-tile vertex bt of the tile container, ntiles x (ny+1) x (nx+1). */
int ko_get_tile_bt(ko_sim* s, double* btt) {
    gather13(s, s->cont[C_TC], NULL, btt);
    return 0;
}

/* CalculateMorphodynamicRHS on a copy of the current state: E - D at cells
 * (global layout) and the vertex bt rate per tile. */
int ko_morpho_rhs(ko_sim* s, double* emd, double* dtbt) {
    init_stepping_arrays(s);
    tdata* I0 = s->cont[C_I0];
    morpho_rhs(s, I0);
    size_t N = (size_t)s->NX * s->NY;
    memset(emd, 0, 8 * N);
    memset(dtbt, 0, 8 * (size_t)s->ntiles * s->NV);
    for (int q = 0; q < s->nactive; ++q) {
        int k = s->active[q];
        scatter_field(s, k, I0[k].emd, emd);
        memcpy(dtbt + (size_t)k * s->NV, I0[k].dtbt, 8 * s->NV);
    }
    return 0;
}

void ko_probe_enable(ko_sim* s, int on) {
    s->probe_on = on;
    s->probe_done = 0;
}

int ko_probe_count(ko_sim* s) { return s->probe_done ? s->probe_n : -1; }

/* The morphodynamic dt of the probed step. */
double ko_probe_dt(ko_sim* s) { return s->probe_dt; }

/* which: 0 = intermed0, 1 = intermed3 before and 2 = after redistribution. */
int ko_probe_get(ko_sim* s, int which, double* f13, double* btt) {
    if (!s->probe_done || which < 0 || which > 2) return KST_ERR_STATE;
    gather13(s, s->probe[which], f13, btt);
    return 0;
}

/* Candidates in the order the check loop marked them (1-based tile ids and
 * cell indices, as AddToRedistList receives them). */
int ko_probe_candidates(ko_sim* s, int32_t* k, int32_t* i, int32_t* j, double* ex, uint8_t* ts) {
    if (!s->probe_done) return KST_ERR_STATE;
    memcpy(ts, s->probe_ts, s->ntiles);
    const rnode* L = (const rnode*)s->probe_list;
    for (int q = 0; q < s->probe_n; ++q) {
        k[q] = L[q].k + 1; i[q] = L[q].i + 1; j[q] = L[q].j + 1; ex[q] = L[q].excess;
    }
    return 0;
}

static void load13(ko_sim* s, tdata* C, const double* f13, const double* btt) {
    size_t N = (size_t)s->NX * s->NY;
    int nx = s->nx, ny = s->ny;
    for (int k = 0; k < s->ntiles; ++k) {
        tdata* T = &C[k];
        if (!T->alloc) continue;
        int gi, gj;
        tile_coords(s, k, &gi, &gj);
        for (int d = 0; d < NF; ++d)
            for (int jj = 0; jj < ny; ++jj)
                for (int ii = 0; ii < nx; ++ii)
                    UF(s, T, d, ii, jj) = f13[d * N + (size_t)(gj * ny + jj) * s->NX + gi * nx + ii];
        if (btt) memcpy(T->bt, btt + (size_t)k * s->NV, 8 * s->NV);
    }
}

/* Overwrite tileContainer's 13 fields (global layout) and, if given, the
 * per-tile vertex bt of every allocated tile, as given. */
int ko_set_fields13(ko_sim* s, const double* f13, const double* btt) {
    load13(s, s->cont[C_TC], f13, btt);
    return 0;
}

/* AddToRedistList over the candidates (1-based tile, i, j, as AddToRedistList
 * receives them) and RedistributeGrid, with intermed0 and intermed3 loaded
 * from (f0, bt0) and (f3, bt3); returns intermed3 and whether it failed. */
int ko_redistribute(ko_sim* s, const double* f0, const double* bt0, const double* f3,
                    const double* bt3, int n, const int32_t* k, const int32_t* i,
                    const int32_t* j, const double* ex, double* fout, double* btout, int* failed) {
    init_stepping_arrays(s);
    load13(s, s->cont[C_I0], f0, bt0);
    load13(s, s->cont[C_I3], f3, bt3);
    rnode* list = NULL;
    int nl = 0, cap = 0;
    for (int q = 0; q < n; ++q) redist_add(&list, &nl, &cap, k[q] - 1, i[q] - 1, j[q] - 1, ex[q]);
    *failed = (nl > 0) ? !redistribute_grid(s, list, nl) : 0;
    free(list);
    gather13(s, s->cont[C_I3], fout, btout);
    return 0;
}

/* Per-tile vertex bt views of every allocated tile (ntiles x (ny+1) x
 * (nx+1)), then the cell topography of those tiles and the face topography
 * of the active ones. */
/* Record (t, dt) of the next `cap` accepted steps into buf (cap = 0 stops);
 * returns the number recorded since the last call that set a buffer. */
int64_t ko_step_log(ko_sim* s, double* buf, int64_t cap) {
    const int64_t n = s->steplog_n;
    s->steplog = cap > 0 ? buf : NULL;
    s->steplog_cap = cap;
    s->steplog_n = 0;
    return n;
}

/* One limiter evaluation, lim(a, b) of src/Limiters.f90:83-184, for the
 * closed-form tests. */
double ko_limiter(int kind, double a, double b) {
    kst_config c;
    memset(&c, 0, sizeof c);
    c.limiter = kind;
    return limiter(&c, a, b);
}

int ko_set_tile_bt(ko_sim* s, const double* btt) {
    for (int k = 0; k < s->ntiles; ++k) {
        tdata* T = &s->cont[C_TC][k];
        if (!T->alloc) continue;
        memcpy(T->bt, btt + (size_t)k * s->NV, 8 * s->NV);
        cell_topo(s, T);
    }
    for (int q = 0; q < s->nactive; ++q) face_topo(s, s->cont[C_TC], s->active[q]);
    return 0;
}
