// CH4-CO2 induced dipole and correlation-function rows on gfx950.
//
// The dipole of the reference (PES-IDS/ai_ids_ch4_co2.cpp:672-1066, PES-IDS/ai_ids_ch4_co2_lib.cpp:
// 8-34) is a sum over 320 products of a radial function and an angular basis function of the
// directions a (intermolecular vector) and b (CO2 axis) in the frame of CH4; expanded
// (ch4co2_dipole_tables.h) it is
//     mu_c = sum_b r_b(R) sum_t coef_t X_{x_t}(a) Y_{y_t}(b),      31802 terms,
// X = P_l^m(a_z) {1, cos m phi1, sin m phi1} (l <= 13, 196 values), Y the same at b for even
// l <= 12 (91 values).  It is too large to share the trajectory kernel's registers, so the
// trajectory kernel records the lab coordinates of every sample (record mode 3) and two kernels
// here finish a job:
//   ch4co2_dipole_kernel  one sample per lane, 64 consecutive samples of one trajectory per
//                         workgroup of kDipWaves wavefronts: the frame of CH4 in every wavefront,
//                         X and Y into LDS (a column per sample; the wavefronts share the orders
//                         m), then each wavefront runs its share of the basis functions (split
//                         by term count) over the term list, read by the whole wavefront at once
//                         (uniform indices: scalar loads), with the radial functions of the
//                         lane's interval (vector loads; the lanes share few intervals, so they
//                         hit the same lines); the shares' partial dipoles are summed in
//                         wavefront order and mu_lab = S1^T mu overwrites the sample's first
//                         three coordinates.  The 287 harmonics of 64 samples take 73 KB (FP32)
//                         or 147 KB (FP64) of LDS, so one wavefront per workgroup would leave
//                         two or three of a CU's four SIMDs idle.
//   ch4co2_cf_rows_kernel one wavefront per trajectory: c_k = mu(0).mu(t_k), the sample checks of
//                         correlation_eval (src/hawaii.c:2181-2236) in sample order, the CF rows
//                         the reduction reads; a failed check marks the trajectory.
// X and Y come from the unit vectors without angles: P_l^m(x) e^{i m phi} =
// (-1)^m d^m P_l/dx^m (x) (a_x + i a_y)^m (Condon-Shortley phase), the derivative by the
// three-term recurrence in l at fixed m.
#include <hip/hip_runtime.h>

#include "ch4co2_dipole_tables.h"
#include "hh_engine.h"
#include "hh_systems.h"

namespace hh {

constexpr int kDipWaves = 4;   // wavefronts per dipole workgroup

template <typename P>
struct Ch4Co2DipTables {
    P radial[HH_CH4CO2_NINT + 1][HH_CH4CO2_NB][4];   // row NINT: the tail (c4, c5, c6, 0)
    P coef[HH_CH4CO2_NTERMS];
    unsigned xy[HH_CH4CO2_NTERMS];   // x | y << 8 (32-bit: read with scalar loads)
    int start[HH_CH4CO2_NB][4];
    double bounds[HH_CH4CO2_NINT];
    double centre[HH_CH4CO2_NINT];
    int bsplit[kDipWaves + 1];   // wavefront w takes basis functions [bsplit[w], bsplit[w + 1])
};

template <typename P>
void fill_ch4co2_tables(Ch4Co2DipTables<P>& t)
{
    for (int k = 0; k < HH_CH4CO2_NINT; ++k) {
        t.bounds[k] = HH_CH4CO2_BOUNDS[k];
        t.centre[k] = HH_CH4CO2_CENTRE[k];
        for (int b = 0; b < HH_CH4CO2_NB; ++b)
            for (int j = 0; j < 4; ++j) t.radial[k][b][j] = (P)HH_CH4CO2_RADIAL[k][b][j];
    }
    for (int b = 0; b < HH_CH4CO2_NB; ++b) {
        for (int j = 0; j < 3; ++j) t.radial[HH_CH4CO2_NINT][b][j] = (P)HH_CH4CO2_TAIL[b][j];
        t.radial[HH_CH4CO2_NINT][b][3] = P(0);
        for (int c = 0; c < 4; ++c) t.start[b][c] = HH_CH4CO2_START[b][c];
    }
    for (int i = 0; i < HH_CH4CO2_NTERMS; ++i) {
        t.coef[i] = (P)HH_CH4CO2_COEF[i];
        t.xy[i] = HH_CH4CO2_XY[i];
    }
    // contiguous basis-function ranges of about equal term counts
    t.bsplit[0] = 0;
    int b = 0;
    for (int w = 1; w < kDipWaves; ++w) {
        const long long goal = (long long)HH_CH4CO2_NTERMS * w / kDipWaves;
        while (b < HH_CH4CO2_NB && HH_CH4CO2_START[b][3] <= goal) ++b;
        t.bsplit[w] = b;
    }
    t.bsplit[kDipWaves] = HH_CH4CO2_NB;
}

size_t ch4co2_table_bytes(bool f32) { return f32 ? sizeof(Ch4Co2DipTables<float>) : sizeof(Ch4Co2DipTables<double>); }

void ch4co2_tables_host(bool f32, void* out)
{
    if (f32)
        fill_ch4co2_tables(*static_cast<Ch4Co2DipTables<float>*>(out));
    else
        fill_ch4co2_tables(*static_cast<Ch4Co2DipTables<double>*>(out));
}

// P_l^m {1, cos, sin} of direction v for l <= LMAX (only even l when EVEN) into column `lane` of
// out[index][64]: index l^2 (m = 0), l^2 + 2m - 1 (cos), l^2 + 2m (sin), or l(l - 1)/2 + ... for
// the even-l table
template <typename P, int LMAX, bool EVEN>
__device__ __forceinline__ void harmonics(const P* v, P (*out)[64], int lane, int wave)
{
    const P x = v[2];
    // w^m with w = -(v_x + i v_y): carries (-1)^m sin^m(theta) e^{i m phi}
    P wr = P(1), wi = P(0);
    const P ur = -v[0], ui = -v[1];
    P dfact = P(1);   // (2m - 1)!!
#pragma unroll
    for (int m = 0; m <= LMAX; ++m) {
        if (m > 0) {
            const P t = wr * ur - wi * ui;
            wi = wr * ui + wi * ur;
            wr = t;
            dfact *= P(2 * m - 1);
        }
        if (m % kDipWaves != wave) continue;   // this order belongs to another wavefront
        // Q_l^m = d^m P_l / dx^m: Q_m^m = (2m - 1)!!, Q_{m+1}^m = (2m + 1) x Q_m^m,
        // (l - m) Q_l^m = (2l - 1) x Q_{l-1}^m - (l + m - 1) Q_{l-2}^m
        P q0 = P(0), q1 = dfact;
#pragma unroll
        for (int l = m; l <= LMAX; ++l) {
            if (l > m) {
                const P q2 = (P(2 * l - 1) * x * q1 - P(l + m - 1) * q0) * P(1.0 / (l - m));
                q0 = q1;
                q1 = q2;
            }
            if (!EVEN || (l % 2) == 0) {
                const int base = EVEN ? l * (l - 1) / 2 : l * l;
                if (m == 0) {
                    out[base][lane] = q1;
                } else {
                    out[base + 2 * m - 1][lane] = q1 * wr;
                    out[base + 2 * m][lane] = q1 * wi;
                }
            }
        }
    }
}

// One workgroup = 64 consecutive samples of one trajectory (blockIdx.y), kDipWaves wavefronts;
// coordinates rec[8] per sample (Phi, Theta, R, phi1, theta1, psi1, phi2, theta2) are replaced by
// (mu_x, mu_y, mu_z).
template <typename P>
__global__ void __launch_bounds__(64 * kDipWaves) ch4co2_dipole_kernel(JobDesc jb, const Ch4Co2DipTables<P>* __restrict__ tb)
{
    __shared__ P sX[HH_CH4CO2_NX][64];
    __shared__ P sY[HH_CH4CO2_NY][64];
    __shared__ P sMu[kDipWaves][3][64];
    const long long item = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int len = jb.len[item];
    const int k0 = blockIdx.x * 64;
    if (k0 >= len) return;                        // uniform: the whole workgroup
    const int k = k0 + lane < len ? k0 + lane : len - 1;   // lanes past the end repeat the last sample
    double* rec = jb.rows + (item * jb.L + k) * 8;
    double q[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) q[i] = rec[i];
    // the frame of CH4 (angles reduced in FP64 by sincos_p), in every wavefront
    P sP, cP, sT, cT, sf, cf, s1, c1, sp, cp, sP2, cP2, s2, c2;
    sincos_p<P>(q[0], &sP, &cP);
    sincos_p<P>(q[1], &sT, &cT);
    sincos_p<P>(q[3], &sf, &cf);
    sincos_p<P>(q[4], &s1, &c1);
    sincos_p<P>(q[5], &sp, &cp);
    sincos_p<P>(q[6], &sP2, &cP2);
    sincos_p<P>(q[7], &s2, &c2);
    const P Rh[3] = {-sT * cP, -sT * sP, cT}, u[3] = {-s2 * cP2, -s2 * sP2, c2};
    P a[3], b[3];
    {
        const P* v[2] = {Rh, u};
        P* o[2] = {a, b};
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const P w0 = cf * v[j][0] + sf * v[j][1], w1 = -sf * v[j][0] + cf * v[j][1], w2 = v[j][2];
            const P x1 = c1 * w1 + s1 * w2, x2 = -s1 * w1 + c1 * w2;
            o[j][0] = cp * w0 + sp * x1;
            o[j][1] = -sp * w0 + cp * x1;
            o[j][2] = x2;
        }
    }
    // each wavefront fills the orders m of its share, for the 64 samples
    harmonics<P, 13, false>(a, sX, lane, wave);
    harmonics<P, 12, true>(b, sY, lane, wave);
    __syncthreads();
    // the lane's radial interval (NINT: the tail beyond the last bound)
    const double R = q[2];
    int kint = 0;
#pragma unroll
    for (int j = 0; j < HH_CH4CO2_NINT; ++j) kint += R >= tb->bounds[j] ? 1 : 0;
    const P t = (P)(R - (kint < HH_CH4CO2_NINT ? tb->centre[kint] : 0.0));
    const P ir = (P)(1.0 / R), ir4 = ir * ir * ir * ir;
    const bool tail = kint == HH_CH4CO2_NINT;
    P mu[3] = {P(0), P(0), P(0)};
    const int b0 = tb->bsplit[wave], b1 = tb->bsplit[wave + 1];
#pragma unroll 1
    for (int bf = b0; bf < b1; ++bf) {
        const P* rc = tb->radial[kint][bf];
        const P r = tail ? (rc[0] + (rc[1] + rc[2] * ir) * ir) * ir4 : ((rc[3] * t + rc[2]) * t + rc[1]) * t + rc[0];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int t0 = tb->start[bf][c], t1 = tb->start[bf][c + 1];
            P s = P(0);
#pragma unroll 8
            for (int e = t0; e < t1; ++e) {
                const unsigned xy = tb->xy[e];
                s += tb->coef[e] * (sX[xy & 0xffu][lane] * sY[xy >> 8][lane]);
            }
            mu[c] += s * r;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) sMu[wave][c][lane] = mu[c];
    __syncthreads();
    if (wave != 0) return;
    // the shares in wavefront order
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        P m = sMu[0][c][lane];
#pragma unroll
        for (int w = 1; w < kDipWaves; ++w) m += sMu[w][c][lane];
        mu[c] = m;
    }
    // mu_lab = Sz(phi1)^T Sx(theta1)^T Sz(psi1)^T mu
    const P x0 = cp * mu[0] - sp * mu[1], x1 = sp * mu[0] + cp * mu[1], x2 = mu[2];
    const P w1 = c1 * x1 - s1 * x2, w2 = s1 * x1 + c1 * x2;
    if (k0 + lane < len) {
        rec[0] = (double)(cf * x0 - sf * w1);
        rec[1] = (double)(sf * x0 + cf * w1);
        rec[2] = (double)w2;
    }
}

// One wavefront per trajectory: c_k = mu(0) . mu(t_k) (FP64 products), the checks of
// correlation_eval in sample order (a NaN dipole; |c| > 1e100; |c_k / c_{k-1}| > 1e10 from k = 2),
// rows cf[item][k] for the samples before the first failing one; a failure sets the item's status
// (the trajectory then counts as zero).
__global__ void __launch_bounds__(64) ch4co2_cf_rows_kernel(JobDesc jb)
{
    const long long item = blockIdx.x;
    const int lane = threadIdx.x;
    const int len = jb.len[item];
    const double* rec = jb.rows + item * jb.L * 8;
    double* cf = jb.cf_rows + item * jb.L;
    const double m0[3] = {rec[0], rec[1], rec[2]};
    double carry = 0.0;
    int fail_k = 0x7fffffff, fail_st = ITEM_OK;
    for (int base = 0; base < len; base += 64) {
        const int k = base + lane;
        const bool valid = k < len;
        double c = 0.0;
        int st = ITEM_OK;
        if (valid) {
            const double* r = rec + (long long)k * 8;
            const double m[3] = {r[0], r[1], r[2]};
            c = m0[0] * m[0] + m0[1] * m[1] + m0[2] * m[2];
            if (isnan(m[0]) || isnan(m[1]) || isnan(m[2])) st = ITEM_FAIL_DIPOLE_NAN;
            else if (fabs(c) > 1e100) st = ITEM_FAIL_CF_HUGE;
        }
        const double up = __shfl_up(c, 1, 64);
        const double prev = lane > 0 ? up : carry;
        carry = __shfl(c, 63, 64);
        if (valid && st == ITEM_OK && k > 1 && fabs(c / prev) > 1e10) st = ITEM_FAIL_CF_JUMP;
        const unsigned long long bad = __ballot(valid && st != ITEM_OK);
        const int first = bad ? __builtin_ctzll(bad) : 64;
        if (valid && lane < first) cf[k] = c;
        if (bad) {
            fail_k = base + first;
            fail_st = __shfl(st, first, 64);
            break;
        }
    }
    if (lane == 0 && fail_k != 0x7fffffff) {
        // the sample checks come before any later step failure of the same trajectory
        jb.status[item] = fail_st;
        jb.len[item] = fail_k;
    }
}

// phase points y (SoA [16][n]) -> the eight coordinates per point as one record row
__global__ void __launch_bounds__(256) ch4co2_records_kernel(const double* __restrict__ y, long long n,
                                                             double* __restrict__ rec, int* __restrict__ len)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) len[0] = (int)n;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 8; ++c) rec[i * 8 + c] = y[(long long)(2 * c) * n + i];
}

__global__ void __launch_bounds__(256) ch4co2_dipole_out_kernel(const double* __restrict__ rec, long long n,
                                                                double* __restrict__ out)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[(long long)c * n + i] = rec[i * 8 + c];
}

hipError_t run_ch4co2_eval_dipole(const double* y, long long n, double* records, int* len, const void* tables,
                                  double* out, hipStream_t st)
{
    const unsigned g = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(ch4co2_records_kernel, dim3(g), dim3(256), 0, st, y, n, records, len);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the points as one item of n samples
    JobDesc jd{};
    jd.rows = records;
    jd.len = len;
    jd.n_items = 1;
    jd.L = (int)n;
    hipLaunchKernelGGL(ch4co2_dipole_kernel<double>, dim3((unsigned)((n + 63) / 64), 1), dim3(64 * kDipWaves), 0, st, jd,
                       static_cast<const Ch4Co2DipTables<double>*>(tables));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(ch4co2_dipole_out_kernel, dim3(g), dim3(256), 0, st, records, n, out);
    return hipGetLastError();
}

hipError_t run_ch4co2_cf_rows(const JobDesc& jd, const void* tables, bool f32, hipStream_t st)
{
    const dim3 grid((unsigned)((jd.L + 63) / 64), (unsigned)jd.n_items);
    if (f32)
        hipLaunchKernelGGL(ch4co2_dipole_kernel<float>, grid, dim3(64 * kDipWaves), 0, st, jd,
                           static_cast<const Ch4Co2DipTables<float>*>(tables));
    else
        hipLaunchKernelGGL(ch4co2_dipole_kernel<double>, grid, dim3(64 * kDipWaves), 0, st, jd,
                           static_cast<const Ch4Co2DipTables<double>*>(tables));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ch4co2_cf_rows_kernel, dim3((unsigned)jd.n_items), dim3(64), 0, st, jd);
    return hipGetLastError();
}

}  // namespace hh
