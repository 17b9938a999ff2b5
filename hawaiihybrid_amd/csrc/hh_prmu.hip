// PR/mu spectral function (calculate_spectral_function_using_prmu_representation_and_save,
// src/hawaii.c:3148-3977) for pairs without requantised rotors: trajectories injected at R0
// with an inward momentum (hh_sample_ics, HH_IC_R0_INJECT), integrated by the batched engine
// with the dipole recorded every sampling_time until the first sample beyond R0 (rows of
// record mode 2), then per trajectory
//     S_t(nu_k) = SF_COEFF * pr_mu * sum_c |FFT(Connes(mu_c))_k|^2,   pr_mu = -pR(0) / m,
// SF_COEFF = Hartree ALU^6 ATU R0^2 dt^2 (src/hawaii.c:3395), summed over the trajectories.
//
// The reference's series starts at t = dt (step 0 of its loop, src/hawaii.c:3578) and holds L
// samples, zero after the trajectory's end; the engine's row also holds mu(0) at index 0, so
// sample i of the series is row index i + 1, and the engine runs L + 2 samples: a row shorter
// than L + 2 ended beyond R0 within the reference's L samples, a row of L + 2 did not (the
// reference drops such a trajectory and draws the next one, src/hawaii.c:3701-3708).
//
// On the device the three components of a trajectory go through two complex FFTs of length L:
// z1 = x + i y gives |X_k|^2 + |Y_k|^2 = (|Z1_k|^2 + |Z1_{L-k}|^2) / 2, z2 = z gives |Z_k|^2
// (the sign convention of the transform does not enter a power spectrum).  The FFT is the
// four-step (Bailey) factorisation L = L1 L2, both factors resident in LDS: with n = n1 L2 + n2
// and k = k1 + L1 k2,
//     y[k1][n2] = w_L^(n2 k1) FFT_L1(x[. L2 + n2])[k1],   X[k1 + L1 k2] = FFT_L2(y[k1][.])[k2],
// one launch for the columns (tiles of adjacent columns, so the strided reads stay in 256 B
// segments) and one for the rows, which writes the spectrum back in natural order (runs of
// adjacent k1 from a padded LDS tile), two passes over HBM instead of log2(L).  Twiddles come
// from a table e^{2 pi i j/L} built once per L on the host in long double, not from a sin/cos
// per butterfly.  The per-trajectory spectra are weighted and summed in trajectory order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/hawaii_gpu.h"
#include "hh_internal.h"

namespace hh {
namespace {

// series s = 2t (x + i y) and 2t + 1 (z) of trajectory t of the chunk, Connes-apodised
// (connes_apodization, src/hawaii.c:5401-5410, over the whole L-sample array)
__global__ void __launch_bounds__(256) prmu_pack_kernel(const double* __restrict__ rows, const int* __restrict__ len,
                                                        const int* __restrict__ keep, long long t0, long long nt,
                                                        int Lrow, long long L, double dt, double2* __restrict__ z)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nt * L) return;
    const long long tl = g / L, i = g % L, t = t0 + tl;
    double x = 0.0, y = 0.0, w = 0.0;
    if (keep[t] && i + 1 < (long long)len[t]) {
        const double* row = rows + t * 3 * (long long)Lrow;
        const double tmax = dt * (double)L, tc = dt * (double)i, r = tc / tmax;
        const double f = (1.0 - r * r) * (1.0 - r * r);
        x = row[i + 1] * f;
        y = row[Lrow + i + 1] * f;
        w = row[2 * (long long)Lrow + i + 1] * f;
    }
    z[(2 * tl) * L + i] = make_double2(x, y);
    z[(2 * tl + 1) * L + i] = make_double2(w, 0.0);
}

constexpr int kFftTile = 4096;   // complex values per workgroup tile: 64 KiB of LDS (plus row padding)

// In place in LDS: `cnt` transforms of length n (a power of two dividing L), X_k = sum_j x_j
// e^{+2 pi i jk/n}: bit reversal, then radix-2 decimation-in-time butterflies, the twiddle
// e^{i pi k/half} read from tw[j] = e^{2 pi i j/L} at j = k L/(2 half) (the table is L2-resident).
// INTERLEAVED: element e of sequence c at a[e * cnt + c], consecutive threads on consecutive
// sequences (the column tiles); else at a[c * stride + e], consecutive threads on consecutive
// butterflies of one sequence (rows; stride = n + 1 keeps a column of the tile, read when it is
// written out, on distinct banks).  Either way consecutive threads touch consecutive LDS words.
__device__ __forceinline__ int ilog2(long long v)
{
    return 63 - __builtin_clzll((unsigned long long)v);
}

template <bool INTERLEAVED>
__device__ void lds_fft_many(double2* a, int n, int cnt, int stride, const double2* __restrict__ tw, long long L,
                             double2* twn)
{
    // every size is a power of two: indices by shifts and masks
    const int logn = ilog2(n), logc = ilog2(cnt), logL = ilog2(L);
    // the n-point transform's twiddles e^{2 pi i m/n}, m < n/2, staged in LDS
    for (int m = threadIdx.x; m < (n >> 1); m += blockDim.x) twn[m] = tw[(long long)m << (logL - logn)];
    auto at = [&](int c, int e) -> double2& { return INTERLEAVED ? a[(e << logc) + c] : a[c * stride + e]; };
    for (int i = threadIdx.x; i < n * cnt; i += blockDim.x) {
        const int c = INTERLEAVED ? i & (cnt - 1) : i >> logn, e = INTERLEAVED ? i >> logc : i & (n - 1);
        const int r = (int)(__builtin_bitreverse32((unsigned)e) >> (32 - logn));
        if (r > e) {
            const double2 t = at(c, e);
            at(c, e) = at(c, r);
            at(c, r) = t;
        }
    }
    __syncthreads();
    const int loghn = logn - 1;
    for (int logh = 0; logh < logn; ++logh) {
        const int half = 1 << logh;
        const int shift = logn - 1 - logh;   // twiddle e^{i pi k/half} = twn[k << shift]
        for (int j = threadIdx.x; j < (n >> 1) * cnt; j += blockDim.x) {
            const int c = INTERLEAVED ? j & (cnt - 1) : j >> loghn, jj = INTERLEAVED ? j >> logc : j & ((n >> 1) - 1);
            const int k = jj & (half - 1);
            const int e0 = ((jj >> logh) << (logh + 1)) + k;
            double2& p0 = at(c, e0);
            double2& p1 = at(c, e0 + half);
            const double2 t = twn[k << shift];
            const double cs = t.x, sn = t.y;
            const double2 u = p0, v = p1;
            const double2 w = make_double2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
            p0 = make_double2(u.x + w.x, u.y + w.y);
            p1 = make_double2(u.x - w.x, u.y - w.y);
        }
        __syncthreads();
    }
}

// step 1: C adjacent columns n2 of one series per workgroup; out[k1 L2 + n2] = twiddled FFT_L1
__global__ void __launch_bounds__(256) prmu_fft_cols_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                            int L1, int L2, int C, long long L,
                                                            const double2* __restrict__ tw)
{
    extern __shared__ double2 tile[];   // [L1][C], then L1/2 twiddles
    const int logc = ilog2(C);
    const int tiles = L2 >> logc;
    const long long s = blockIdx.x / tiles;
    const int n2_0 = (int)(blockIdx.x % tiles) << logc;
    const double2* x = in + s * L;
    for (int i = threadIdx.x; i < C * L1; i += blockDim.x) {
        const int n1 = i >> logc, c = i & (C - 1);
        tile[i] = x[(long long)n1 * L2 + n2_0 + c];
    }
    __syncthreads();
    lds_fft_many<true>(tile, L1, C, 0, tw, L, tile + C * L1);
    double2* y = out + s * L;
    for (int i = threadIdx.x; i < C * L1; i += blockDim.x) {
        const int k1 = i >> logc, n2 = n2_0 + (i & (C - 1));
        const double2 v = tile[i];
        const double2 t = tw[((long long)n2 * k1) & (L - 1)];   // e^{2 pi i n2 k1 / L}
        y[(long long)k1 * L2 + n2] = make_double2(v.x * t.x - v.y * t.y, v.x * t.y + v.y * t.x);
    }
}

// step 2: R adjacent rows k1 of one series per workgroup, FFT_L2 of each; the result is written
// in natural order, X[k1 + L1 k2] at out[k1 + L1 k2] (runs of R consecutive k1)
__global__ void __launch_bounds__(256) prmu_fft_rows_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                            int L1, int L2, int R, long long L,
                                                            const double2* __restrict__ tw)
{
    extern __shared__ double2 tile[];   // [R][L2 + 1], then L2/2 twiddles
    const int logr = ilog2(R), log2n = ilog2(L2);
    const int tiles = L1 >> logr;
    const int stride = L2 + 1;
    const long long s = blockIdx.x / tiles;
    const int k1_0 = (int)(blockIdx.x % tiles) << logr;
    const double2* x = in + s * L + (long long)k1_0 * L2;
    for (int i = threadIdx.x; i < R * L2; i += blockDim.x) tile[(i >> log2n) * stride + (i & (L2 - 1))] = x[i];
    __syncthreads();
    lds_fft_many<false>(tile, L2, R, stride, tw, L, tile + R * stride);
    double2* y = out + s * L;
    for (int i = threadIdx.x; i < R * L2; i += blockDim.x) {
        const int r = i & (R - 1), k2 = i >> logr;
        y[k1_0 + r + (long long)L1 * k2] = tile[r * stride + k2];
    }
}

// pw[t][k] = w[t] * (|X_k|^2 + |Y_k|^2 + |Z_k|^2), k < nfreq (gsl_fft_square, src/hawaii.c:2349-2356)
__global__ void __launch_bounds__(256) prmu_power_kernel(const double2* __restrict__ Z, long long L,
                                                         int nfreq, long long t0, long long nt,
                                                         const double* __restrict__ w, double* __restrict__ pw)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nt * nfreq) return;
    const long long tl = g / nfreq, k = g % nfreq, km = (L - k) & (L - 1);
    const double2* z1 = Z + (2 * tl) * L;
    const double2* z2 = Z + (2 * tl + 1) * L;
    const double2 a = z1[k], b = z1[km], c = z2[k];
    const double p = 0.5 * ((a.x * a.x + a.y * a.y) + (b.x * b.x + b.y * b.y)) + (c.x * c.x + c.y * c.y);
    pw[tl * nfreq + k] = w[t0 + tl] * p;
}

// sf[k] += sum over the chunk's trajectories, in trajectory order
__global__ void __launch_bounds__(256) prmu_accumulate_kernel(const double* __restrict__ pw, long long nt, int nfreq,
                                                              double* __restrict__ sf)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nfreq) return;
    double acc = sf[k];
    for (long long t = 0; t < nt; ++t) acc += pw[t * nfreq + k];
    sf[k] = acc;
}

constexpr double kSfHartree = 4.3597447222071e-18;   // src/constants.h:10
constexpr double kSfALU = 5.29177210903e-11;         // src/constants.h:8
constexpr double kSfATU = 2.4188843265857e-17;       // src/constants.h:12

}  // namespace
}  // namespace hh

using hh::fail;

extern "C" int hh_prmu(hh_engine* e, const hh_system* sys, const hh_prmu_params* p, const double* ic_dev, long ntraj,
                       double* sf_dev, int* counted, long* ncounted, hh_cf_stats* stats)
{
    using namespace hh;
    if (!e || !sys || !p || !ic_dev || !sf_dev || ntraj <= 0) return fail(HH_E_INVALID, "hh_prmu: bad argument");
    if (p->L < 4 || p->L > (1 << 20) || (p->L & (p->L - 1)))
        return fail(HH_E_INVALID, "hh_prmu: MaxTrajectoryLength must be a power of two in [4, 2^20] (radix-2 FFT)");
    if (p->nfreq < 1 || p->nfreq > p->L / 2 + 1)
        return fail(HH_E_INVALID, "hh_prmu: nfreq must lie in [1, L/2 + 1]");
    if (!(p->dt > 0) || !(p->R0 > 0) || !(p->rtol > 0) || !(p->atol > 0))
        return fail(HH_E_INVALID, "hh_prmu: sampling_time, R0 and tolerances must be positive");
    hh_cf_params cp{};
    cp.rtol = p->rtol;
    cp.atol = p->atol;
    cp.dt = p->dt;
    cp.Rcut = p->R0;
    cp.ppf = 1.0;
    cp.L = p->L + 2;
    cp.halves = HH_HALVES_FORWARD;
    cp.max_steps = p->max_steps;
    cp.engine = HH_ENGINE_BATCHED;
    cp.precision = p->precision;
    cp.orbit_rtol = p->orbit_rtol;
    double* dummy[1] = {sf_dev};   // no CF is reduced for a series job
    const double* icp[1] = {ic_dev};
    const long nt[1] = {ntraj};
    // held across the run and the read of its rows, so no other call on the engine replaces them
    std::lock_guard<std::recursive_mutex> lock(e->mu);
    int rc = cf_run(e, sys, &cp, 1, icp, nt, dummy, stats, 2);
    if (rc < 0) return rc;
    HH_CHECK(hipSetDevice(e->device));
    const JobDesc jd = e->last_jobs[0];
    std::vector<int> len(ntraj), status(ntraj), keep(ntraj);
    std::vector<double> pR(ntraj), w(ntraj);
    HH_CHECK(hipMemcpy(len.data(), jd.len, ntraj * sizeof(int), hipMemcpyDeviceToHost));
    HH_CHECK(hipMemcpy(status.data(), jd.status, ntraj * sizeof(int), hipMemcpyDeviceToHost));
    HH_CHECK(hipMemcpy(pR.data(), ic_dev + 5 * ntraj, ntraj * sizeof(double), hipMemcpyDeviceToHost));
    const double SF_COEFF = kSfHartree * std::pow(kSfALU, 6) * kSfATU * p->R0 * p->R0 * p->dt * p->dt;
    long n_keep = 0;
    for (long t = 0; t < ntraj; ++t) {
        // ended beyond R0 within L samples (or kept truncated), without a failed dipole or step
        keep[t] = status[t] == 0 && (len[t] <= p->L + 1 || p->allow_truncating) ? 1 : 0;
        const double pr_mu = -pR[t] / sys->mu;
        w[t] = keep[t] ? SF_COEFF * pr_mu : 0.0;
        n_keep += keep[t];
        if (counted) counted[t] = keep[t];
    }
    if (ncounted) *ncounted = n_keep;
    const long long L = p->L;
    int logL = 0;
    while ((1LL << logL) < L) ++logL;
    const int L1 = 1 << ((logL + 1) / 2), L2 = 1 << (logL / 2);   // L = L1 L2, both <= 1024
    const int C = std::min(kFftTile / L1, L2), R = std::min(kFftTile / L2, L1);
    // chunks of trajectories whose two series per trajectory fit 1 GiB per buffer
    const long long chunk = std::max<long long>(1, std::min<long long>(ntraj, (1LL << 30) / (2 * L * 16)));
    if (e->prmu_tw_L != L) {
        // e^{2 pi i j/L}, rounded from long double
        std::vector<double2> tw((size_t)L);
        for (long long j = 0; j < L; ++j) {
            const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)L;
            tw[(size_t)j] = make_double2((double)cosl(a), (double)sinl(a));
        }
        e->prmu_tw_L = 0;
        HH_CHECK(e->prmu_tw.ensure((size_t)L * sizeof(double2)));
        HH_CHECK(hipMemcpyAsync(e->prmu_tw.p, tw.data(), (size_t)L * sizeof(double2), hipMemcpyHostToDevice, e->stream));
        HH_CHECK(hipStreamSynchronize(e->stream));
        e->prmu_tw_L = L;
    }
    const double2* twd = e->prmu_tw.as<double2>();
    const size_t cols_lds = ((size_t)C * L1 + L1 / 2) * sizeof(double2);
    const size_t rows_lds = ((size_t)R * (L2 + 1) + L2 / 2) * sizeof(double2);
    HH_CHECK(hipFuncSetAttribute((const void*)prmu_fft_cols_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)cols_lds));
    HH_CHECK(hipFuncSetAttribute((const void*)prmu_fft_rows_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)rows_lds));
    HH_CHECK(e->prmu_a.ensure((size_t)chunk * 2 * L * sizeof(double2)));
    HH_CHECK(e->prmu_b.ensure((size_t)chunk * 2 * L * sizeof(double2)));
    HH_CHECK(e->prmu_w.ensure((size_t)ntraj * (sizeof(double) + sizeof(int)) + sizeof(int) +
                              (size_t)chunk * p->nfreq * sizeof(double)));
    double* wd = e->prmu_w.as<double>();
    int* keepd = reinterpret_cast<int*>(wd + ntraj);
    double* pw = reinterpret_cast<double*>(keepd + ntraj + (ntraj & 1));
    HH_CHECK(hipMemcpyAsync(wd, w.data(), ntraj * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HH_CHECK(hipMemcpyAsync(keepd, keep.data(), ntraj * sizeof(int), hipMemcpyHostToDevice, e->stream));
    HH_CHECK(hipEventRecord(e->ev[0], e->stream));
    for (long long t0 = 0; t0 < ntraj; t0 += chunk) {
        const long long n = std::min<long long>(chunk, ntraj - t0);
        double2* a = e->prmu_a.as<double2>();
        double2* b = e->prmu_b.as<double2>();
        hipLaunchKernelGGL(prmu_pack_kernel, dim3((unsigned)((n * L + 255) / 256)), dim3(256), 0, e->stream, jd.rows,
                           jd.len, keepd, t0, n, cp.L, L, p->dt, a);
        HH_CHECK(hipGetLastError());
        // four-step transform a -> b (columns) -> a (rows)
        hipLaunchKernelGGL(prmu_fft_cols_kernel, dim3((unsigned)(2 * n * (L2 / C))), dim3(256), cols_lds,
                           e->stream, a, b, L1, L2, C, L, twd);
        HH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(prmu_fft_rows_kernel, dim3((unsigned)(2 * n * (L1 / R))), dim3(256), rows_lds, e->stream,
                           b, a, L1, L2, R, L, twd);
        HH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(prmu_power_kernel, dim3((unsigned)((n * p->nfreq + 255) / 256)), dim3(256), 0, e->stream, a,
                           L, p->nfreq, t0, n, wd, pw);
        HH_CHECK(hipGetLastError());
        hipLaunchKernelGGL(prmu_accumulate_kernel, dim3((unsigned)((p->nfreq + 255) / 256)), dim3(256), 0, e->stream, pw,
                           n, p->nfreq, sf_dev);
        HH_CHECK(hipGetLastError());
    }
    HH_CHECK(hipEventRecord(e->ev[1], e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    if (stats) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[1]);
        stats->reduce_ms += ms;
    }
    return HH_OK;
}
