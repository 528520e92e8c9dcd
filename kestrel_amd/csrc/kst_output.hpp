// kst_output.hpp -- device side of the output diagnostics.
//
// CalculateVolume (src/Output.f90:617-735) sums over the active tiles' cells
// the flow volume (w - b0 - bt) gam^2, the mass rho Hn gam, the bed change bt
// and the solids Hnpsi gam (Hn and rho as ComputeDesingularisedVariables,
// src/HydraulicRHS.f90:762-878, derives them from the stored state).  The
// reference Kahan-sums them sequentially in active-list order.  Here each sum
// is a compensated parallel sum, Sum2 of Ogita, Rump and Oishi ("Accurate
// sum and dot product", SIAM J. Sci. Comput. 26, 2005): every addition is
// an error-free TwoSum (Knuth) whose rounding error is accumulated beside
// the running sum, partial (sum, error) pairs merge the same way, and the
// result is sum + error.  Its error is at most u|S| + gamma_{n-1}^2 sum|x_i|,
// so for these sums both it and the reference's Kahan sum lie within an ulp
// or two of the exact sum; the order of a launch is fixed by the active set,
// so the result is deterministic.
//
// The pass reads five FP64 words per active cell (w, Hnpsi, b0, bt, gam):
// an HBM-bound stream, one tile per workgroup iteration, coalesced along the
// tile's rows.
#pragma once

namespace kst {

constexpr int VOL_NT = 256;            // threads of a volume workgroup
constexpr int VOL_BLOCKS_MAX = 2048;   // partials of the first pass (8 XCDs x 256)
constexpr int VOL_Q = 4;               // flow volume, mass, bed, solids (in this order)

struct Sum2 {
    double s = 0.0, e = 0.0;
    __device__ __forceinline__ void add(double x) {
        const double t = s + x;
        const double z = t - s;
        e = e + ((s - (t - z)) + (x - z));
        s = t;
    }
    __device__ __forceinline__ void merge(double s2, double e2) {
        add(s2);
        e = e + e2;
    }
};

// Merges the VOL_Q pairs of a workgroup into thread 0 (wave shuffles, then
// the waves' pairs through LDS).
__device__ __forceinline__ void vol_block_reduce(Sum2 (&a)[VOL_Q], double* sh) {
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < VOL_Q; ++q) {
            const double s2 = __shfl_xor(a[q].s, off, 64);
            const double e2 = __shfl_xor(a[q].e, off, 64);
            a[q].merge(s2, e2);
        }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < VOL_Q; ++q) {
            sh[(wv * VOL_Q + q) * 2] = a[q].s;
            sh[(wv * VOL_Q + q) * 2 + 1] = a[q].e;
        }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < VOL_NT / 64; ++w)
            for (int q = 0; q < VOL_Q; ++q) a[q].merge(sh[(w * VOL_Q + q) * 2], sh[(w * VOL_Q + q) * 2 + 1]);
}

// First pass: workgroup b takes active-list entries b, b + G, ... and writes
// its VOL_Q (sum, error) pairs to part[b].
__global__ __launch_bounds__(VOL_NT) void volume_kernel(const Dev* dA, const Params* pA, double* part) {
    const auto& d = const_arg(dA);
    const auto& p = const_arg(pA);
    __shared__ double sh[VOL_NT / 64 * VOL_Q * 2];
    const auto* c = ctrl_view(d);
    const int na = c->n_active;
    // mixed precision: u^n is in S0 or S0b by Ctrl::cur
    const double* Sn = (d.mixed && c->cur) ? d.S0b : d.S0;
    const long long FS = d.FS;
    const int NC = p.NC;
    const double eps2 = p.hthr * p.hthr;
    Sum2 a[VOL_Q];
    for (int it = blockIdx.x; it < na; it += gridDim.x) {
        const long long base = (long long)d.active[it] * NC;
        for (int i = threadIdx.x; i < NC; i += VOL_NT) {
            const long long g = base + i;
            const double w = Sn[g], hp = Sn[3 * FS + g];
            const double b0 = d.b0c[g], bt = d.TA.btc[g], gam = d.TA.cgam[g];
            // (w - b0 - bt) gam gam, evaluated left to right
            double dv = w - b0;
            dv = dv - bt;
            dv = dv * gam;
            dv = dv * gam;
            double hn = compute_hn(w, b0, bt, gam);
            if (hn < 0.0) hn = 0.0;
            const double hpc = hp < 0.0 ? 0.0 : hp;
            const double psi = dmin(2.0 * hn * hpc / (hn * hn + dmax(hn * hn, eps2)), p.max_pack);
            const double rho = density(generic(p), psi);
            a[0].add(dv);
            a[1].add(rho * hn * gam);
            a[2].add(bt);
            a[3].add(hp * gam);
        }
    }
    vol_block_reduce(a, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < VOL_Q; ++q) {
            part[(size_t)blockIdx.x * VOL_Q * 2 + 2 * q] = a[q].s;
            part[(size_t)blockIdx.x * VOL_Q * 2 + 2 * q + 1] = a[q].e;
        }
}

// Second pass: one workgroup merges the nb partials (thread t takes t,
// t + VOL_NT, ...) and writes the VOL_Q pairs to out.
__global__ __launch_bounds__(VOL_NT) void volume_final_kernel(const double* part, int nb, double* out) {
    __shared__ double sh[VOL_NT / 64 * VOL_Q * 2];
    Sum2 a[VOL_Q];
    for (int b = threadIdx.x; b < nb; b += VOL_NT)
        for (int q = 0; q < VOL_Q; ++q)
            a[q].merge(part[(size_t)b * VOL_Q * 2 + 2 * q], part[(size_t)b * VOL_Q * 2 + 2 * q + 1]);
    vol_block_reduce(a, sh);
    if (threadIdx.x == 0)
        for (int q = 0; q < VOL_Q; ++q) {
            out[2 * q] = a[q].s;
            out[2 * q + 1] = a[q].e;
        }
}

// ---------------------------------------------------------------------------
// OutputSolutionData_txt (src/Output.f90:754-929): per active cell, in the
// order of the file's columns 6-19, flow depth Hn, flow speed, u, v, rho,
// psi, rhoHnu, rhoHnv, Hnpsi, w, b0, bt, db/dx, db/dy, as tileContainer holds
// them.  Hn, psi and rho are its desingularised variables
// (ComputeDesingularisedVariables, src/HydraulicRHS.f90:762-878, evaluated as
// the reference orders it), which the final implicit pass leaves consistent
// with w and Hnpsi; u and v are the ones the step's last RHS derived (Dev::UVo),
// since that pass changes the momentum after them.  The speed is
// sqrt(FlowSquaredSpeedSlopeAligned) of those where Hn exceeds the height
// threshold, else 0 (src/Output.f90:868-873).  out[list position][column][cell].
// ---------------------------------------------------------------------------

constexpr int OUT_COLS = 14;

// tileContainer's velocities after kst_set_state: desingularised from the
// state (the reference's set-up ends with ComputeDesingularisedVariables on
// the active tiles).
__global__ __launch_bounds__(VOL_NT) void uv_init_kernel(const Dev* dA, const Params* pA) {
    const auto& d = const_arg(dA);
    const auto& p = const_arg(pA);
    const auto* c = ctrl_view(d);
    const int na = c->n_active;
    const double* Sn = (d.mixed && c->cur) ? d.S0b : d.S0;
    const long long FS = d.FS;
    const int NC = p.NC;
    const double eps = p.hthr;
    for (int it = blockIdx.x; it < na; it += gridDim.x) {
        const long long base = (long long)d.active[it] * NC;
        for (int i = threadIdx.x; i < NC; i += VOL_NT) {
            const long long g = base + i;
            const double w = Sn[g], ru = Sn[FS + g], rv = Sn[2 * FS + g], hp = Sn[3 * FS + g];
            double hn = compute_hn(w, d.b0c[g], d.TA.btc[g], d.TA.cgam[g]);
            if (hn < 0.0) hn = 0.0;
            const double hpc = hp < 0.0 ? 0.0 : hp;
            const double psi = dmin(2.0 * hn * hpc / (hn * hn + dmax(hn * hn, eps * eps)), p.max_pack);
            const double rho = density(generic(p), psi);
            d.UVo[g] = 2.0 * hn * ru / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
            // a 1D grid never derives v (src/HydraulicRHS.f90:870-874)
            d.UVo[FS + g] = p.oned ? 0.0 : 2.0 * hn * rv / (hn * hn + dmax(hn * hn, eps * eps)) / rho;
        }
    }
}

__global__ __launch_bounds__(VOL_NT) void output_kernel(const Dev* dA, const Params* pA, double* out) {
    const auto& d = const_arg(dA);
    const auto& p = const_arg(pA);
    const auto* c = ctrl_view(d);
    const int na = c->n_active;
    const double* Sn = (d.mixed && c->cur) ? d.S0b : d.S0;
    const long long FS = d.FS;
    const int NC = p.NC;
    const double eps = p.hthr;
    for (int it = blockIdx.x; it < na; it += gridDim.x) {
        const long long base = (long long)d.active[it] * NC;
        double* o = out + (size_t)it * OUT_COLS * NC;
        for (int i = threadIdx.x; i < NC; i += VOL_NT) {
            const long long g = base + i;
            const double w = Sn[g], ru = Sn[FS + g], rv = Sn[2 * FS + g], hp = Sn[3 * FS + g];
            const double b0 = d.b0c[g], bt = d.TA.btc[g], bx = d.TA.cbx[g], by = d.TA.cby[g];
            const double gam = d.TA.cgam[g];
            double hn = compute_hn(w, b0, bt, gam);
            if (hn < 0.0) hn = 0.0;
            const double hpc = hp < 0.0 ? 0.0 : hp;
            const double psi = dmin(2.0 * hn * hpc / (hn * hn + dmax(hn * hn, eps * eps)), p.max_pack);
            const double rho = density(generic(p), psi);
            const double uu = d.UVo[g], vv = d.UVo[FS + g];
            const double spd = hn > p.hthr ? sqrt(speed2(p.geometric, uu, vv, bx, by)) : 0.0;
            const double col[OUT_COLS] = {hn, spd, uu, vv, rho, psi, ru, rv, hp, w, b0, bt, bx, by};
#pragma unroll
            for (int q = 0; q < OUT_COLS; ++q) o[(size_t)q * NC + i] = col[q];
        }
    }
}

}  // namespace kst
