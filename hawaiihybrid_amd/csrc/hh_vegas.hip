// Adaptive Monte Carlo (VEGAS) phase-space integrals on the device.
//
// Integrands (src/hep_hawaii.cpp:116-201): a point x of [0,1]^QP maps to phase space by
// transform_variables (src/hep_hawaii.cpp:8-114): R = 1/x0 (Jacobian R^2), every momentum
// p = tan(pi (x - 1/2)) (Jacobian pi (1 + p^2)), Phi = 2 pi x0, Theta = pi x2, linear-molecule
// angles phi = 2 pi x6, theta = pi x8.  The reference reads x0 for both R and Phi and never
// reads x4 (it indexes the point with the phase-space slot of Phi, IPHI = 0); this is kept, so
// the same x gives the same phase point (the integral is unaffected: Phi only enters through
// the relative orientation, which phi1 covers uniformly).  Outside [Rmin, Rmax], or off the pair
// state (FREE: E >= 0, BOUND: E <= 0), the integrand is zero; else
//     PF: J exp(-E HkT / T),      M0: J |mu|^2 exp(-E HkT / T),
//     M2: J |d mu/dt|^2 exp(-E HkT / T) (integrand_M2, src/hep_hawaii.cpp:228-330; the dipole's
//         coordinate derivatives by central differences, h = 1e-3, contracted with dH/dp).
//
// The integrator follows G. P. Lepage, J. Comput. Phys. 27 (1978) 192: a separable grid of
// kVegasBins bins per dimension, each iteration samples N points uniformly in grid space
// (Philox4x32-10 counter-based uniforms, J. K. Salmon et al., SC'11: point i of iteration it is
// counter (i, it), key (seed)), estimates I_it = <f J_grid> and its variance, and refines the
// grid from the per-bin sums of (f J_grid)^2 (smoothing over neighbours, compression
// ((1 - r)/ln(1/r))^alpha with alpha = 1.5, equal-weight redistribution).  The result is the
// inverse-variance weighted mean of every iteration but the first, as mpi_perform_integration
// accumulates hep-mc's iterations (src/hep_hawaii.cpp:348-355).  hep-mc 0.7 itself is not
// available: the estimate agrees with the reference's statistically, not bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/hawaii_gpu.h"
#include "hh_internal.h"

namespace hh {

constexpr int kVegasBins = 128;
constexpr int kVegasBlock = 256;
constexpr int kVegasMaxDim = 10;
constexpr double kVegasHkT = 4.3597447222071e-18 / 1.380649e-23;   // src/constants.h:24
constexpr double kVegasPI = 3.14159265358979323846;

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

struct VegasParams {
    double T, Rmin, Rmax;
    int pair_state;
    int integrand;          // HH_INTEGRAND_*
    int dim;                // QP
    int iteration;
    uint32_t seed;
    long long npoints;
};

// M2 selects the M2 integrand at compile time, so the M0 and PF kernels keep their registers
template <int SYS, bool M2>
__device__ __forceinline__ double vegas_integrand(const SysConst& sc, const VegasParams& p, const double* x)
{
    constexpr int QP = SysTraits<SYS>::QP;
    double y[QP];
    double jac = 1.0;
    // transform_variables: slot IR reads x[0], slot IPR reads x[5]
    y[4] = 1.0 / x[0];
    jac *= y[4] * y[4];
    y[5] = tan(kVegasPI * (x[5] - 0.5));
    jac *= kVegasPI * (1.0 + y[5] * y[5]);
    y[0] = x[0] * 2.0 * kVegasPI;
    jac *= 2.0 * kVegasPI;
    y[1] = tan(kVegasPI * (x[1] - 0.5));
    jac *= kVegasPI * (1.0 + y[1] * y[1]);
    y[2] = x[2] * kVegasPI;
    jac *= kVegasPI;
    y[3] = tan(kVegasPI * (x[3] - 0.5));
    jac *= kVegasPI * (1.0 + y[3] * y[3]);
    if constexpr (SYS == SYS_CO2AR) {
        y[6] = x[6] * 2.0 * kVegasPI;
        jac *= 2.0 * kVegasPI;
        y[7] = tan(kVegasPI * (x[7] - 0.5));
        jac *= kVegasPI * (1.0 + y[7] * y[7]);
        y[8] = x[8] * kVegasPI;
        jac *= kVegasPI;
        y[9] = tan(kVegasPI * (x[9] - 0.5));
        jac *= kVegasPI * (1.0 + y[9] * y[9]);
    }
    const double R = y[4];
    if (R < p.Rmin || R > p.Rmax) return 0.0;
    const double energy = kinetic<SYS>(sc, y) + pes_value<SYS, double>(sc, y);
    if (p.pair_state == HH_PAIR_FREE && energy < 0.0) return 0.0;
    if (p.pair_state == HH_PAIR_BOUND && energy > 0.0) return 0.0;
    double w = jac * exp(-energy * kVegasHkT / p.T);
    if constexpr (M2) {
        w *= dipole_rate_sq<SYS>(sc, y);
    } else if (p.integrand == HH_INTEGRAND_M0) {
        double mu[3];
        dipole<SYS, double>(sc, y, mu);
        w *= mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2];
    }
    return w;
}

// Every block writes sum f, sum f^2 (f = integrand x grid Jacobian), the count of non-finite
// values and, per dimension and bin, sum f^2 of its points.
template <int SYS, bool M2>
__global__ __launch_bounds__(kVegasBlock) void vegas_kernel(SysConst sc, VegasParams p, const double* __restrict__ grid,
                                                            double* __restrict__ block_sums,
                                                            double* __restrict__ block_bins)
{
    constexpr int D = SysTraits<SYS>::QP;
    __shared__ double edges[D][kVegasBins + 1];
    __shared__ double bins[D][kVegasBins];
    __shared__ double red[3][kVegasBlock / 64];
    for (int i = threadIdx.x; i < D * (kVegasBins + 1); i += blockDim.x) edges[i / (kVegasBins + 1)][i % (kVegasBins + 1)] = grid[i];
    for (int i = threadIdx.x; i < D * kVegasBins; i += blockDim.x) bins[i / kVegasBins][i % kVegasBins] = 0.0;
    __syncthreads();
    double s1 = 0.0, s2 = 0.0, nnf = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < p.npoints; i += stride) {
        double x[kVegasMaxDim];
        int b[kVegasMaxDim];
        double jac = 1.0;
#pragma unroll
        for (int g = 0; g < (D + 3) / 4; ++g) {
            uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)p.iteration, (uint32_t)g};
            philox10(c, p.seed, 0x48415741u);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int d = 4 * g + j;
                if (d >= D) break;
                const double u = ((double)c[j] + 0.5) * (1.0 / 4294967296.0);   // in (0, 1)
                const double yb = u * kVegasBins;
                int bi = (int)yb;
                bi = bi < kVegasBins ? bi : kVegasBins - 1;
                const double lo = edges[d][bi], w = edges[d][bi + 1] - lo;
                x[d] = lo + (yb - bi) * w;
                jac *= kVegasBins * w;
                b[d] = bi;
            }
        }
        double f = vegas_integrand<SYS, M2>(sc, p, x) * jac;
        if (!isfinite(f)) {
            f = 0.0;
            nnf += 1.0;
        }
        s1 += f;
        s2 += f * f;
        if (f != 0.0) {
#pragma unroll
            for (int d = 0; d < D; ++d) atomicAdd(&bins[d][b[d]], f * f);
        }
    }
    // block sums: wavefront reduction, then the waves in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s1 += __shfl_down(s1, off);
        s2 += __shfl_down(s2, off);
        nnf += __shfl_down(nnf, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = s1;
        red[1][threadIdx.x >> 6] = s2;
        red[2][threadIdx.x >> 6] = nnf;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double a = 0.0;
        for (int w = 0; w < kVegasBlock / 64; ++w) a += red[threadIdx.x][w];
        block_sums[(long long)blockIdx.x * 3 + threadIdx.x] = a;
    }
    for (int i = threadIdx.x; i < D * kVegasBins; i += blockDim.x)
        block_bins[(long long)blockIdx.x * D * kVegasBins + i] = bins[i / kVegasBins][i % kVegasBins];
}

// sums over blocks in block order (one thread per output value)
__global__ __launch_bounds__(256) void vegas_sum_kernel(const double* __restrict__ in, int nblocks, int width,
                                                        double* __restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= width) return;
    double s = 0.0, c = 0.0;
    for (int b = 0; b < nblocks; ++b) {
        const double v = in[(long long)b * width + j];
        const double t = s + v;
        c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
        s = t;
    }
    out[j] = s + c;
}

namespace {

// Lepage's refinement of one dimension's bin edges from its per-bin sums of f^2
void refine(double* edges, const double* d_in, double alpha)
{
    const int n = kVegasBins;
    double d[kVegasBins], r[kVegasBins];
    double total = 0.0;
    for (int i = 0; i < n; ++i) total += d_in[i];
    if (!(total > 0.0)) return;
    // smoothing with the neighbours
    d[0] = (d_in[0] + d_in[1]) / 2.0;
    d[n - 1] = (d_in[n - 2] + d_in[n - 1]) / 2.0;
    for (int i = 1; i < n - 1; ++i) d[i] = (d_in[i - 1] + d_in[i] + d_in[i + 1]) / 3.0;
    double sd = 0.0;
    for (int i = 0; i < n; ++i) sd += d[i];
    double rsum = 0.0;
    for (int i = 0; i < n; ++i) {
        const double q = d[i] / sd;
        r[i] = (q > 0.0 && q < 1.0) ? std::pow((q - 1.0) / std::log(q), alpha) : (q >= 1.0 ? 1.0 : 0.0);
        rsum += r[i];
    }
    if (!(rsum > 0.0)) return;
    // new edges: each new bin holds rsum / n of the weight, the weight spread evenly inside old bins
    double old[kVegasBins + 1];
    std::memcpy(old, edges, sizeof(old));
    const double per = rsum / n;
    double acc = 0.0;
    int k = 0;
    for (int i = 1; i < n; ++i) {
        const double target = per * i;
        while (k < n && acc + r[k] < target) acc += r[k++];
        if (k >= n) {
            edges[i] = old[n];
            continue;
        }
        const double frac = r[k] > 0.0 ? (target - acc) / r[k] : 0.0;
        edges[i] = old[k] + frac * (old[k + 1] - old[k]);
    }
    edges[0] = 0.0;
    edges[n] = 1.0;
}

}  // namespace
}  // namespace hh

extern "C" int hh_vegas(hh_engine* e, const hh_system* sys, const hh_vegas_params* vp, hh_vegas_result* out)
{
    using namespace hh;
    if (!e || !sys || !vp || !out) return fail(HH_E_INVALID, "hh_vegas: null argument");
    if (sys->system != HH_SYS_HEAR && sys->system != HH_SYS_CO2AR)
        return fail(HH_E_UNSUPPORTED, "hh_vegas: system %d has no device surfaces", sys->system);
    if (vp->integrand != HH_INTEGRAND_M0 && vp->integrand != HH_INTEGRAND_M2 && vp->integrand != HH_INTEGRAND_PF)
        return fail(HH_E_INVALID, "hh_vegas: unknown integrand %d", vp->integrand);
    if (!(vp->T > 0) || !(vp->Rmin > 0) || !(vp->Rmax > vp->Rmin))
        return fail(HH_E_INVALID, "hh_vegas: T > 0 and 0 < sampler_Rmin < sampler_Rmax required");
    if (vp->niterations < 2 || vp->npoints < 2)
        return fail(HH_E_INVALID, "hh_vegas: at least 2 iterations of 2 points (the first is discarded)");
    const int D = hh_qp_size(sys->system);
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    HH_CHECK(hipSetDevice(e->device));
    const int nblocks = (int)std::min<long long>(4096, (vp->npoints + kVegasBlock - 1) / kVegasBlock);
    const size_t width_bins = (size_t)D * kVegasBins;
    DevBuf gridb, sums, binsb, outb;
    HH_CHECK(gridb.ensure((size_t)D * (kVegasBins + 1) * sizeof(double)));
    HH_CHECK(sums.ensure((size_t)nblocks * 3 * sizeof(double)));
    HH_CHECK(binsb.ensure((size_t)nblocks * width_bins * sizeof(double)));
    HH_CHECK(outb.ensure((3 + width_bins) * sizeof(double)));
    std::vector<double> grid((size_t)D * (kVegasBins + 1));
    for (int d = 0; d < D; ++d)
        for (int i = 0; i <= kVegasBins; ++i) grid[(size_t)d * (kVegasBins + 1) + i] = (double)i / kVegasBins;
    SysConst sc{sys->mu, sys->I, e->tab64.p};
    VegasParams p{vp->T, vp->Rmin, vp->Rmax, vp->pair_state, vp->integrand, D, 0, vp->seed, vp->npoints};
    std::vector<double> host(3 + width_bins);
    double wsum = 0.0, wisum = 0.0, chi_a = 0.0;   // sums of weights, weighted values, weighted squares
    long long nnf_total = 0;
    int used = 0;
    HH_CHECK(hipEventRecord(e->ev[0], e->stream));
    for (int it = 0; it < vp->niterations; ++it) {
        p.iteration = it;
        HH_CHECK(hipMemcpyAsync(gridb.p, grid.data(), grid.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
        const bool m2 = vp->integrand == HH_INTEGRAND_M2;
        auto k = sys->system == HH_SYS_HEAR ? (m2 ? vegas_kernel<SYS_HEAR, true> : vegas_kernel<SYS_HEAR, false>)
                                            : (m2 ? vegas_kernel<SYS_CO2AR, true> : vegas_kernel<SYS_CO2AR, false>);
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(kVegasBlock), 0, e->stream, sc, p, gridb.as<double>(),
                           sums.as<double>(), binsb.as<double>());
        HH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(vegas_sum_kernel, dim3(1), dim3(256), 0, e->stream, sums.as<double>(), nblocks, 3,
                           outb.as<double>());
        hipLaunchKernelGGL(vegas_sum_kernel, dim3((unsigned)((width_bins + 255) / 256)), dim3(256), 0, e->stream,
                           binsb.as<double>(), nblocks, (int)width_bins, outb.as<double>() + 3);
        HH_CHECK(hipGetLastError());
        HH_CHECK(hipMemcpyAsync(host.data(), outb.p, host.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HH_CHECK(hipStreamSynchronize(e->stream));
        const double N = (double)vp->npoints;
        const double mean = host[0] / N;
        const double var = std::max(0.0, (host[1] / N - mean * mean) / (N - 1.0));
        nnf_total += (long long)host[2];
        if (it > 0 && var > 0.0) {
            const double w = 1.0 / var;
            wsum += w;
            wisum += w * mean;
            chi_a += w * mean * mean;
            ++used;
        }
        for (int d = 0; d < D; ++d) refine(&grid[(size_t)d * (kVegasBins + 1)], &host[3 + (size_t)d * kVegasBins], 1.5);
        out->iter_value[it < 32 ? it : 31] = mean;
        out->iter_error[it < 32 ? it : 31] = std::sqrt(var);
    }
    HH_CHECK(hipEventRecord(e->ev[1], e->stream));
    HH_CHECK(hipEventSynchronize(e->ev[1]));
    float t = 0.0f;
    HH_CHECK(hipEventElapsedTime(&t, e->ev[0], e->ev[1]));
    if (used == 0 || !(wsum > 0.0)) return fail(HH_E_INVALID, "hh_vegas: the integrand vanished on every point");
    out->value = wisum / wsum;
    out->error = 1.0 / std::sqrt(wsum);
    out->chi2_dof = used > 1 ? (chi_a - wisum * wisum / wsum) / (used - 1) : 0.0;
    out->iterations = vp->niterations;
    out->points = (long long)vp->npoints * vp->niterations;
    out->nonfinite = nnf_total;
    out->ms = t;
    return HH_OK;
}
