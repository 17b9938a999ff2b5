// Zimmermann's trick: the CF of a trajectory averaged over every time origin of its dipole
// series (correlation_eval_zimmerman_trick, src/hawaii.c:1888-2114).
//
// The reference fills a series d[0 .. 2L-2] per trajectory (backward half reversed into
// d[0 .. L-1], mu(0) at d[L-1], forward half in d[L-1 .. 2L-2], zero after a half stops) and
// sums, for every shift s < L, over the first L origins:
//     crln[s] = ALU^3 / L * sum_{c=0}^{L-1} d(c) . d(c+s)          (O(L^2) per trajectory)
// Here that sum is a circular cross-correlation of a = d[0 .. L-1] (zero-padded) with d on
// P points (the power of two >= 2L - 1), exact because c + s <= 2L - 2 < P.  With F the transform
// X[k] = sum_j x[j] e^{+2 pi i jk/P}:
//     sum_c a[c] d[c+s] = (1/P) sum_k conj(A[k]) D[k] e^{-2 pi i ks/P}.
// Per trajectory and component one complex FFT of z = a + i d gives both spectra
// (A = (Z[k] + conj Z[P-k])/2, D = (Z[k] - conj Z[P-k])/2i).  The ensemble needs only the sum
// of conj(A) D over trajectories and components, a Hermitian spectrum accumulated per bin
// k = 0..P/2 in a fixed order (workgroup w takes trajectories w, w + G, ...; Neumaier sums), so
// a job's CF does not depend on what else ran.  One inverse transform per job finishes it.
// Each transform runs in LDS (P complex doubles: 128 KiB at L = 4096).
#include <hip/hip_runtime.h>

#include <cmath>

#include "hh_engine.h"
#include "hh_systems.h"

namespace hh {

constexpr int kZimmGroups = 256;   // reduction workgroups (fixed: it fixes the summation order)
constexpr int kZimmThreads = 1024;
constexpr int kZimmMaxL = 4096;
constexpr int kZimmBins = (kZimmMaxL + 1 + kZimmThreads - 1) / kZimmThreads;   // bins per thread

__device__ __forceinline__ void zimm_neumaier(double& s, double& c, double v)
{
    const double t = s + v;
    if (fabs(s) >= fabs(v))
        c += (s - t) + v;
    else
        c += (v - t) + s;
    s = t;
}

// In-place radix-2 transform X[k] = sum_j x[j] e^{+2 pi i jk/P} of P = 2^logP points in LDS;
// tw[j] = e^{+2 pi i j/P}, j < P/2.  Ends with a barrier.
__device__ void lds_fft(double2* a, int P, int logP, const double2* __restrict__ tw)
{
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int r = (int)(__brev((unsigned)i) >> (32 - logP));
        if (r > i) {
            const double2 t = a[i];
            a[i] = a[r];
            a[r] = t;
        }
    }
    __syncthreads();
    for (int half = 1, step = P / 2; half < P; half <<= 1, step >>= 1) {
        for (int j = threadIdx.x; j < P / 2; j += blockDim.x) {
            const int k = j & (half - 1);
            const int i0 = ((j - k) << 1) + k, i1 = i0 + half;
            const double2 w = tw[k * step];
            const double2 u = a[i0], v = a[i1];
            const double2 t = make_double2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
            a[i0] = make_double2(u.x + t.x, u.y + t.y);
            a[i1] = make_double2(u.x - t.x, u.y - t.y);
        }
        __syncthreads();
    }
}

// partial[w][k] (re, im; with Neumaier compensations) = sum over this workgroup's trajectories
// and the three components of ppf conj(A[k]) D[k], k = 0..P/2.
__global__ __launch_bounds__(kZimmThreads) void zimm_partial_kernel(JobDesc jb, double ppf, int logP,
                                                                    const double2* tw, double* __restrict__ psum,
                                                                    double* __restrict__ pcomp)
{
    extern __shared__ double2 z[];
    const int L = jb.L, P = 1 << logP, H = P / 2;
    double sr[kZimmBins], si[kZimmBins], cr[kZimmBins], ci[kZimmBins];
#pragma unroll
    for (int m = 0; m < kZimmBins; ++m) sr[m] = si[m] = cr[m] = ci[m] = 0.0;
    for (long long t = blockIdx.x; t < jb.ntraj; t += gridDim.x) {
        const long long fi = 2 * t, bi = 2 * t + 1;   // forward, backward item
        if (dropped(jb.status[fi], jb.status[bi])) continue;
        const int lf = jb.len[fi], lb = jb.len[bi];
        const double* rf = jb.rows + fi * 3LL * L;
        const double* rb = jb.rows + bi * 3LL * L;
        for (int comp = 0; comp < 3; ++comp) {
            for (int j = threadIdx.x; j < P; j += blockDim.x) {
                // series index j: backward sample L-1-j for j < L-1, forward sample j-(L-1) from L-1 on
                double d;
                if (j >= L - 1) {
                    const int k = j - (L - 1);
                    d = k < lf ? rf[comp * L + k] : 0.0;
                } else {
                    const int k = L - 1 - j;
                    d = k < lb ? rb[comp * L + k] : 0.0;
                }
                z[j] = make_double2(j < L ? d : 0.0, d);
            }
            __syncthreads();
            lds_fft(z, P, logP, tw);
#pragma unroll
            for (int m = 0; m < kZimmBins; ++m) {
                const int k = threadIdx.x + m * kZimmThreads;
                if (k > H) continue;
                const double2 Zk = z[k], Zm = z[(P - k) & (P - 1)];
                // A = (Zk + conj Zm)/2, D = (Zk - conj Zm)/(2i)
                const double Ar = 0.5 * (Zk.x + Zm.x), Ai = 0.5 * (Zk.y - Zm.y);
                const double Dr = 0.5 * (Zk.y + Zm.y), Di = -0.5 * (Zk.x - Zm.x);
                // conj(A) D
                zimm_neumaier(sr[m], cr[m], ppf * (Ar * Dr + Ai * Di));
                zimm_neumaier(si[m], ci[m], ppf * (Ar * Di - Ai * Dr));
            }
            __syncthreads();
        }
    }
    const long long base = (long long)blockIdx.x * 2 * (H + 1);
#pragma unroll
    for (int m = 0; m < kZimmBins; ++m) {
        const int k = threadIdx.x + m * kZimmThreads;
        if (k > H) continue;
        psum[base + 2 * k] = sr[m];
        psum[base + 2 * k + 1] = si[m];
        pcomp[base + 2 * k] = cr[m];
        pcomp[base + 2 * k + 1] = ci[m];
    }
}

// cf[s] = ALU^3 / L / ntraj * (1/P) sum_k S[k] e^{-2 pi i ks/P}, S the Hermitian sum over groups
__global__ __launch_bounds__(kZimmThreads) void zimm_final_kernel(int L, int logP, int groups, long long ntraj,
                                                                  const double2* tw, const double* __restrict__ psum,
                                                                  const double* __restrict__ pcomp,
                                                                  double* __restrict__ cf_out)
{
    extern __shared__ double2 y[];
    const int P = 1 << logP, H = P / 2;
    for (int k = threadIdx.x; k <= H; k += blockDim.x) {
        double s_r = 0.0, s_i = 0.0, c_r = 0.0, c_i = 0.0;
        for (int g = 0; g < groups; ++g) {
            const long long o = (long long)g * 2 * (H + 1) + 2 * k;
            zimm_neumaier(s_r, c_r, psum[o]);
            zimm_neumaier(s_i, c_i, psum[o + 1]);
            c_r += pcomp[o];
            c_i += pcomp[o + 1];
        }
        const double Sr = s_r + c_r, Si = s_i + c_i;
        // y = conj(S) on 0..P/2, Hermitian extension S on P-k
        y[k] = make_double2(Sr, -Si);
        if (k > 0 && k < H) y[P - k] = make_double2(Sr, Si);
    }
    __syncthreads();
    lds_fft(y, P, logP, tw);   // Y[s] = conj(P r[s])
    const double scale = kALU * kALU * kALU / (double)L / (double)P / (double)ntraj;
    for (int s = threadIdx.x; s < L; s += blockDim.x) cf_out[s] = y[s].x * scale;
}

int zimm_log_points(int L)
{
    int logP = 1;
    while ((1 << logP) < 2 * L - 1) ++logP;
    return logP;
}

size_t zimm_workspace_doubles(int L) { return (size_t)kZimmGroups * 2 * ((1 << zimm_log_points(L)) / 2 + 1); }

int zimm_max_length() { return kZimmMaxL; }

hipError_t run_zimm_reduce(const JobDesc& jb, double ppf, const double2* tw, double* psum, double* pcomp,
                           double* cf_out, hipStream_t st)
{
    const int logP = zimm_log_points(jb.L);
    const size_t lds = ((size_t)1 << logP) * sizeof(double2);
    // the LDS limit is a per-device attribute
    static unsigned long long attr_set = 0;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (!(attr_set & (1ull << (dev & 63)))) {
        e = hipFuncSetAttribute((const void*)zimm_partial_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(2 * kZimmMaxL * sizeof(double2)));
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute((const void*)zimm_final_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(2 * kZimmMaxL * sizeof(double2)));
        if (e != hipSuccess) return e;
        attr_set |= 1ull << (dev & 63);
    }
    hipLaunchKernelGGL(zimm_partial_kernel, dim3(kZimmGroups), dim3(kZimmThreads), lds, st, jb, ppf, logP, tw, psum,
                       pcomp);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(zimm_final_kernel, dim3(1), dim3(kZimmThreads), lds, st, jb.L, logP, kZimmGroups, jb.ntraj, tw,
                       psum, pcomp, cf_out);
    return hipGetLastError();
}

}  // namespace hh
