// CF -> SF transform and the Monte Carlo M0 moment on gfx950.
//
// CF -> SF (idct_cf_to_sf, src/hawaii.c:5569-5601):
//   Connes apodization in place (src/hawaii.c:5401-5410), Makhoul packing of the IDCT into a
//   2n-point complex sequence (src/hawaii.c:5478-5553), a backward complex FFT, then the real
//   part of the first n outputs times tmax/pi*Yscale at nu_k = k pi/tmax*Xscale.
//   The FFT is radix-2: resident in LDS (one workgroup, one global read and one global write)
//   while 2n complex doubles fit the workgroup's LDS, else one global Stockham pass per stage.
//
// M0 (calculate_M0, src/hawaii.c:1423-1491): every lane owns one Mersenne Twister stream
// (mt_seed32 seeding, state interleaved across lanes so equal draw indices coalesce) and
// runs the reference's candidate loop (q_generator, p_generator, Hamiltonian, reject, pair
// filter, dipole, Welford running mean / sum of squared deviations) over its quota.
// M2 (calculate_M2, src/hawaii.c:1524-1617) is the same loop with |d mu/dt|^2 as the value.
#include <hip/hip_runtime.h>

#include "hh_engine.h"
#include "hh_systems.h"

namespace hh {

constexpr double kPI = 3.14159265358979323846;

// ------------------------------------------------------------------ CF -> SF
// The pack kernel only reads cf (thread k > n needs the original cf[2n - k] that thread 2n - k
// would otherwise be overwriting); the in-place apodization of the input that the reference
// performs is written by sf_unpack_kernel, after the transform.
__global__ void __launch_bounds__(256) sf_pack_kernel(const double* __restrict__ cf, long n, double dt,
                                                      double2* __restrict__ z)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 2 * n) return;
    const double tmax = dt * n;
    const double inv2n = 1.0 / (2.0 * n);
    double s, c;
    sincospi((double)k / (2.0 * n), &s, &c);   // e^{i pi k / 2n}
    double v;
    if (k < n) {
        const double tc = dt * k, r = tc / tmax;
        const double f = (1.0 - r * r) * (1.0 - r * r);
        v = cf[k] * f;
    } else if (k == n) {
        z[k] = make_double2(0.0, 0.0);
        return;
    } else {
        const long i = 2 * n - k;
        const double tc = dt * i, r = tc / tmax;
        v = -(cf[i] * ((1.0 - r * r) * (1.0 - r * r)));
    }
    z[k] = make_double2(v * c * inv2n, v * s * inv2n);
}

// one radix-2 Stockham stage over global memory: len = 2^stage * 2
__global__ void __launch_bounds__(256) fft_stage_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                        long M, long half_len)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;   // butterfly index, M/2 of them
    if (j >= M / 2) return;
    const long k = j % half_len;        // position inside the sub-transform
    const long g = j / half_len;        // sub-transform group
    // Stockham: inputs a = in[j], b = in[j + M/2]; outputs to out[g*2*half_len + k] and + half_len
    const double2 a = in[j], b = in[j + M / 2];
    double s, c;
    sincospi((double)k / (double)half_len, &s, &c);   // backward: w = e^{+i pi k / half_len}
    const double2 wb = make_double2(b.x * c - b.y * s, b.x * s + b.y * c);
    out[g * 2 * half_len + k] = make_double2(a.x + wb.x, a.y + wb.y);
    out[g * 2 * half_len + k + half_len] = make_double2(a.x - wb.x, a.y - wb.y);
}

// whole transform in LDS for M <= 8192 complex values (128 KiB)
__global__ void __launch_bounds__(1024) fft_lds_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                       long M)
{
    extern __shared__ double2 sm[];
    double2* a = sm;
    for (long i = threadIdx.x; i < M; i += blockDim.x) a[i] = in[i];
    __syncthreads();
    // in place: bit reversal, then decimation-in-time radix-2 butterflies
    int logM = 0;
    while ((1L << logM) < M) ++logM;
    for (long i = threadIdx.x; i < M; i += blockDim.x) {
        long r = 0, x = i;
        for (int b = 0; b < logM; ++b) {
            r = (r << 1) | (x & 1);
            x >>= 1;
        }
        if (r > i) {
            const double2 t = a[i];
            a[i] = a[r];
            a[r] = t;
        }
    }
    __syncthreads();
    for (long half = 1; half < M; half <<= 1) {
        for (long j = threadIdx.x; j < M / 2; j += blockDim.x) {
            const long k = j % half;
            const long i0 = (j / half) * 2 * half + k, i1 = i0 + half;
            double s, c;
            sincospi((double)k / (double)half, &s, &c);
            const double2 u = a[i0], v = a[i1];
            const double2 w = make_double2(v.x * c - v.y * s, v.x * s + v.y * c);
            a[i0] = make_double2(u.x + w.x, u.y + w.y);
            a[i1] = make_double2(u.x - w.x, u.y - w.y);
        }
        __syncthreads();
    }
    for (long i = threadIdx.x; i < M; i += blockDim.x) out[i] = a[i];
}

__global__ void __launch_bounds__(256) sf_unpack_kernel(const double2* __restrict__ Z, long n, double dt,
                                                        double* __restrict__ nu, double* __restrict__ sf,
                                                        double* __restrict__ cf)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    {
        // Connes apodization of the input in place, as the reference does (src/hawaii.c:5581)
        const double r = (dt * k) / (dt * n);
        cf[k] *= (1.0 - r * r) * (1.0 - r * r);
    }
    const double LightSpeed_cm = 2.99792458e10, ATU = 2.4188843265857e-17;
    const double ADIPMOMU = 8.4783536255e-30, EPSILON0 = 8.8541878128e-12;
    const double Xscale = 1.0 / LightSpeed_cm / ATU / 2.0 / kPI;
    const double Yscale = ATU * ADIPMOMU * ADIPMOMU / (4.0 * kPI * EPSILON0);
    const double tmax = n * dt;
    nu[k] = k * kPI / tmax * Xscale;
    sf[k] = Z[k].x * tmax / kPI * Yscale;
}

hipError_t run_cf_to_sf(double* cf, long n, double dt, double2* work_a, double2* work_b, double* nu, double* sf,
                        hipStream_t st, bool lds_ok)
{
    const long M = 2 * n;
    hipLaunchKernelGGL(sf_pack_kernel, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, st, cf, n, dt, work_a);
    double2* res = work_a;
    if (M <= 8192 && lds_ok) {
        const size_t lds = M * sizeof(double2);
        if (lds > 65536) {
            hipError_t e = hipFuncSetAttribute((const void*)fft_lds_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(fft_lds_kernel, dim3(1), dim3(1024), lds, st, work_a, work_b, M);
        res = work_b;
    } else {
        double2 *src = work_a, *dst = work_b;
        for (long half = 1; half < M; half <<= 1) {
            hipLaunchKernelGGL(fft_stage_kernel, dim3((unsigned)((M / 2 + 255) / 256)), dim3(256), 0, st, src, dst, M,
                               half);
            double2* t = src;
            src = dst;
            dst = t;
        }
        res = src;
    }
    hipLaunchKernelGGL(sf_unpack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, res, n, dt, nu, sf, cf);
    return hipGetLastError();
}

// ------------------------------------------------------------------ CF -> SF without allocation
// The two 2n-point work sequences of the transform live in a file-scope device array instead of
// engine-allocated buffers, for callers that cannot allocate (graph capture, tight serving loops).
// Every device holds its own copy; its address comes from hipGetSymbolAddress once per device
// (noalloc_scratch_address, at engine creation) and only that address is passed to kernels: the
// array's name in host code is a host-side stand-in, not a device address.  Calls that overlap on
// one device share the array and race.
__device__ __align__(256) double2 scratch[2 * 2 * kCfToSfNoAllocMaxN];

hipError_t noalloc_scratch_address(void** ptr)
{
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

// ------------------------------------------------------------------ M0
constexpr int kMTN = 624, kMTM = 397;
constexpr double kHkT = 4.3597447222071e-18 / 1.380649e-23;

// Stream state for lane s lives at st[i * nstreams + s].
struct LaneMt {
    uint32_t* st;
    long ns, s;
    int pos;
    __device__ uint32_t& w(int i) { return st[(long)i * ns + s]; }
    __device__ void twist()
    {
        for (int i = 0; i < kMTN; ++i) {
            const uint32_t y = (w(i) & 0x80000000u) | (w((i + 1) % kMTN) & 0x7fffffffu);
            w(i) = w((i + kMTM) % kMTN) ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        pos = 0;
    }
    __device__ double drand()
    {
        if (pos == kMTN) twist();
        uint32_t y = w(pos++);
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return (double)y * (1.0 / 4294967296.0);
    }
};

__device__ double m0_normal(LaneMt& mt)
{
    for (;;) {
        const double U = mt.drand();
        const double V = mt.drand();
        const double logU = log(U);
        if (isnan(logU) || isinf(logU)) continue;
        return sqrt(-2 * logU) * cos(2.0 * kPI * V);
    }
}

// One launch draws candidates for every stream until it has its quota, hits the candidate
// limit, or (ticks > 0) the launch has run `ticks` wall-clock ticks; `init` seeds the streams.
// ORDER 0 averages |mu|^2 (M0), ORDER 2 |d mu/dt|^2 (M2).
template <int SYS, int ORDER>
__global__ void __launch_bounds__(256) m0_kernel(SysConst sc, M0Params p, uint32_t seed, long nstreams, long npoints,
                                                 uint32_t* __restrict__ state, M0Acc acc, int init, long long ticks)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nstreams) return;
    const long long t_end = wall_clock64() + ticks;
    LaneMt mt{state, nstreams, s, kMTN};
    const long quota = npoints / nstreams + (s < npoints % nstreams ? 1 : 0);
    const long max_candidates = quota * 100000 + 100000;   // bounded even if nothing is ever accepted
    long candidates = 0, n = 0;
    double m = 0.0, q = 0.0;
    if (init) {
        uint32_t v = seed + (uint32_t)s;
        if (v == 0) v = 4357u;
        mt.w(0) = v;
        for (int i = 1; i < kMTN; ++i) mt.w(i) = 69069u * mt.w(i - 1);
        mt.twist();
    } else {
        mt.pos = acc.pos[s];
        m = acc.m[s];
        q = acc.q[s];
        n = acc.n[s];
        candidates = acc.cand[s];
    }
    const bool was_done = !init && (n >= quota || candidates >= max_candidates);
    double y[10];
    const double RMIN = p.Rmin, RMAX = p.Rmax, T = p.T;
    const double sqrt_MUKT = sqrt(sc.mu / kHkT * T);
    const double sqrt_IIKT = sqrt(sc.I / kHkT * T);
    const double M = 1.5 * exp(-p.pesmin * kHkT / T);
    while (n < quota && candidates < max_candidates) {
        if (ticks > 0 && wall_clock64() >= t_end) break;
        ++candidates;
        const double R3 = mt.drand() * (RMAX * RMAX * RMAX - RMIN * RMIN * RMIN) + RMIN * RMIN * RMIN;
        y[4] = pow(R3, 1.0 / 3.0);
        y[0] = mt.drand() * 2.0 * kPI;
        y[2] = acos(2.0 * mt.drand() - 1.0);
        if (SYS == SYS_CO2AR) {
            y[6] = mt.drand() * 2.0 * kPI;
            y[8] = acos(2.0 * mt.drand() - 1.0);
        }
        const double sinT = sin(y[2]);
        const double x0 = m0_normal(mt), x1 = m0_normal(mt), x2 = m0_normal(mt);
        y[1] = sqrt_MUKT * y[4] * sinT * x0;
        y[3] = sqrt_MUKT * y[4] * x1;
        y[5] = sqrt_MUKT * x2;
        if (SYS == SYS_CO2AR) {
            const double st = sin(y[8]);
            const double r0 = m0_normal(mt), r1 = m0_normal(mt);
            y[7] = sqrt_IIKT * st * r1;
            y[9] = sqrt_IIKT * r0;
        }
        // Hamiltonian before reject (src/hawaii.c:1456-1459), then reject (src/hawaii.c:1382-1421):
        // u is drawn first; K and V at the same point
        const double K = kinetic<SYS>(sc, y);
        const double V = pes_value<SYS, double>(sc, y);
        const double energy = K + V;
        const double u = mt.drand();
        const double weight = exp(-(K + V) * kHkT / T) / exp(-K * kHkT / T) / M;
        if (weight < u) continue;
        if (p.pair_state == 1 && energy < 0.0) continue;
        if (p.pair_state == 2 && energy > 0.0) continue;
        double f;
        if constexpr (ORDER == 2) {
            f = dipole_rate_sq<SYS>(sc, y);
        } else {
            double mu[3];
            dipole<SYS, double>(sc, y, mu);
            f = mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2];
        }
        const double diff = f - m;
        m += diff / (n + 1.0);
        q += diff * diff * (n / (n + 1.0));
        ++n;
    }
    acc.m[s] = m;
    acc.q[s] = q;
    acc.n[s] = n;
    acc.cand[s] = candidates;
    acc.pos[s] = mt.pos;
    if (!was_done && (n >= quota || candidates >= max_candidates)) atomicAdd(acc.finished, 1);
}

hipError_t run_m0(int sys, const SysConst& sc, const M0Params& p, uint32_t seed, long nstreams, long npoints,
                  uint32_t* state, const M0Acc& acc, int init, long long ticks, hipStream_t st, int order)
{
    const unsigned grid = (unsigned)((nstreams + 255) / 256);
    auto k = sys == SYS_HEAR ? (order == 2 ? m0_kernel<SYS_HEAR, 2> : m0_kernel<SYS_HEAR, 0>)
                             : (order == 2 ? m0_kernel<SYS_CO2AR, 2> : m0_kernel<SYS_CO2AR, 0>);
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, st, sc, p, seed, nstreams, npoints, state, acc, init, ticks);
    return hipGetLastError();
}

}  // namespace hh
