// Trajectory-ensemble kernels for gfx950.
//
//  * plain_cf_kernel      one work item (one half of one trajectory) per thread, FP64,
//                         runs its item to the end: the straightforward GPU port of
//                         correlation_eval (src/hawaii.c:2116-2347).
//  * persistent_cf_kernel resident lanes pull items from a device work queue; a lane whose
//                         item ends (R > Rcut, L samples, failure) takes the next item in the
//                         same loop iteration, so uneven trajectory lengths do not idle the
//                         wavefront.  Surface evaluation in P (float in the mixed engine).
//  * cf_reduce_*          ensemble sum over trajectories of 1/2 (c_f + c_b) ALU^3 ppf in a
//                         fixed order with Neumaier-compensated FP64 partial sums, then / ntraj
//                         (src/hawaii.c:2334-2337, 3007-3009, 3111-3114).
//
// Integrator: DOP853 (Dormand-Prince 8(5,3) + 7th-order dense output; Hairer, Norsett &
// Wanner II.5-II.6), error norm sk = atol + rtol max(|y|,|y_new|), controller safe 0.9,
// fac1 0.333, fac2 6, beta 0, HINIT start.  Samples at t_k = k dt from the dense output
// (CVode CV_NORMAL interpolation, src/trajectory.c:155).  State and error control in FP64.
//
// Stage storage: the 16 stage vectors K0..K15 and the 7 dense-output coefficient vectors
// F0..F6 of a lane live in a per-wavefront arena in HBM, interleaved across the 64 lanes,
// so every stage access is a coalesced 64-lane load or store that stays in L2 / MALL.  Only y, the
// stage input and the right-hand-side temporaries occupy registers; holding all stage
// vectors in registers needs ~440 registers per lane for CO2-Ar (one wave per SIMD).
#include <hip/hip_runtime.h>

#include "dop853_coeffs.h"
#include "hh_engine.h"
#include "hh_systems.h"

namespace hh {

constexpr double kEPS = 2.220446049250313e-16;
constexpr int kMaxTotalSteps = 400000;   // attempts per item; the longest CO2-Ar items take ~6e3
constexpr int kSlots = 16;               // K0..K15
#ifndef HH_MIN_WAVES
#define HH_MIN_WAVES 2   // occupancy target (waves per SIMD) of the trajectory kernels
#endif

// Wave-blocked arena: wavefront w owns kSlots*N*64 consecutive doubles; components are paired,
// element (slot, i) of lane l at double offset (slot*N + (i & ~1))*64 + 2l + (i & 1) of the
// block, so one 16-byte access moves components i, i+1 of a lane and a wavefront's access is
// 1 KiB contiguous.  Accesses go through a buffer resource of the block (wave-uniform, scalar
// registers), a 32-bit lane offset and a constant slot offset: no per-lane 64-bit address is
// kept live.  An access outside the block is dropped (stores) or reads 0 (loads).
typedef unsigned int hh_v2u __attribute__((ext_vector_type(2)));
typedef unsigned int hh_v4u __attribute__((ext_vector_type(4)));
constexpr int kBufferConfig = 0x00020000;   // raw 32-bit-element buffer, gfx9 word 3

template <int N>
struct Arena {
    static_assert(N % 2 == 0, "components are stored in pairs");
    __amdgpu_buffer_rsrc_t rsrc;
    int lane16;
    static __device__ __forceinline__ int off(int slot, int i) { return (slot * N + (i & ~1)) * 512 + (i & 1) * 8; }
    __device__ __forceinline__ double ld(int slot, int i) const
    {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, lane16, off(slot, i), 0));
    }
    __device__ __forceinline__ void st(int slot, int i, double v) const
    {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(hh_v2u, v), rsrc, lane16, off(slot, i), 0);
    }
    // components i, i + 1 (i even) in one access
    __device__ __forceinline__ void ld2(int slot, int i, double& a, double& b) const
    {
        const hh_v4u v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, lane16, (slot * N + i) * 512, 0);
        a = __builtin_bit_cast(double, hh_v2u{v.x, v.y});
        b = __builtin_bit_cast(double, hh_v2u{v.z, v.w});
    }
    __device__ __forceinline__ void st2(int slot, int i, double a, double b) const
    {
        const hh_v2u x = __builtin_bit_cast(hh_v2u, a), y = __builtin_bit_cast(hh_v2u, b);
        __builtin_amdgcn_raw_buffer_store_b128(hh_v4u{x.x, x.y, y.x, y.y}, rsrc, lane16, (slot * N + i) * 512, 0);
    }
    // components i, i + 1 (i even) of another lane of the wavefront
    __device__ __forceinline__ void ld2_lane(int lane, int slot, int i, double& a, double& b) const
    {
        const hh_v4u v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, lane * 16, (slot * N + i) * 512, 0);
        a = __builtin_bit_cast(double, hh_v2u{v.x, v.y});
        b = __builtin_bit_cast(double, hh_v2u{v.z, v.w});
    }
};

template <int N>
__device__ __forceinline__ Arena<N> wave_arena(double* ws, long long wave)
{
    const long long w = __builtin_amdgcn_readfirstlane((int)wave);
    double* block = ws + w * (long long)(kSlots * N * 64);
    return Arena<N>{__builtin_amdgcn_make_buffer_rsrc(block, 0, kSlots * N * 64 * 8, kBufferConfig),
                    (int)(threadIdx.x & 63) * 16};
}

template <int N>
struct Lane {
    double y[N];
    double t, h, E0;
    double mu0[3];
    double prev, curr;
    long long local;   // item index inside its job
    int job;
    int k;             // next sample index
    int steps;         // accepted steps since the last recorded sample
    int total_steps;   // step attempts of the item
    int nrhs;
    int reject;
    int active;
};

__device__ __forceinline__ int find_job(const JobDesc* __restrict__ jobs, int njobs, long long g)
{
    int j = 0;
    while (j + 1 < njobs && g >= jobs[j + 1].item_begin) ++j;
    return j;
}

// Hairer's HINIT for an order-8 method (one extra right-hand side); f0 is in arena slot 0
template <int SYS, typename P, int N>
__device__ __forceinline__ double initial_step(const SysConst& sc, const Arena<N>& ar, const double* y0, double rtol,
                                               double atol, double hmax, int& nrhs)
{
    double dnf = 0.0, dny = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double sk = atol + rtol * fabs(y0[i]);
        const double f = ar.ld(0, i);
        dnf += (f / sk) * (f / sk);
        dny += (y0[i] / sk) * (y0[i] / sk);
    }
    double h = (dnf <= 1e-10 || dny <= 1e-10) ? 1e-6 : sqrt(dny / dnf) * 0.01;
    h = fmin(h, hmax);
    double y1[N], f1[N];
#pragma unroll
    for (int i = 0; i < N; ++i) y1[i] = y0[i] + h * ar.ld(0, i);
    rhs<SYS, P>(sc, y1, f1);
    ++nrhs;
    double der2 = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double sk = atol + rtol * fabs(y0[i]);
        const double r = (f1[i] - ar.ld(0, i)) / sk;
        der2 += r * r;
    }
    der2 = sqrt(der2) / h;
    const double der12 = fmax(fabs(der2), sqrt(dnf));
    const double h1 = (der12 <= 1e-15) ? fmax(1e-6, fabs(h) * 1e-3) : sqrt(sqrt(sqrt(0.01 / der12)));
    return fmin(fmin(100.0 * fabs(h), h1), hmax);
}

// `last` is the index of the last stored sample (samples [0, last] are in the row).
template <int N>
__device__ __forceinline__ void finish_item(const JobDesc& jb, Lane<N>& ln, int status, int last, float drift)
{
    jb.len[ln.local] = last + 1;
    jb.status[ln.local] = status;
    // an orbit job adds its right-hand sides to those the mixed-precision run spent on the item
    if (jb.map) {
        jb.nrhs[ln.local] += ln.nrhs;
        atomicAdd(jb.orbit_nrhs, (unsigned long long)ln.nrhs);
    } else
        jb.nrhs[ln.local] = ln.nrhs;
    jb.e_drift[ln.local] = drift;
    if (jb.done) atomicAdd(jb.done, 1);
    ln.active = 0;
}

// Start item `local` of job `j`: load the initial condition (momenta inverted for the
// backward half, invert_momenta src/hawaii.c:1171-1187), mu(0), c[0], H(0), f(y0), first
// step size.  Returns false if the item ended during set-up.
template <int SYS, typename P, int N>
__device__ __forceinline__ bool init_item(const SysConst& sc, const JobDesc& jb, int j, long long local,
                                          const Arena<N>& ar, Lane<N>& ln)
{
    ln.job = j;
    ln.local = jb.map ? (long long)jb.map[local] : local;
    const long long traj = ln.local / jb.halves;
    const bool back = (ln.local % jb.halves) == 1;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double v = jb.ic[(long long)i * jb.ntraj + traj];
        ln.y[i] = (back && (i & 1)) ? -v : v;
    }
    double* row = jb.rows + ln.local * (long long)jb.L * row_width(jb.record);
    if constexpr (SYS == SYS_CH4CO2) {
        // record 3: the coordinates of sample 0; the dipole comes later (hh_ch4co2.hip)
#pragma unroll
        for (int c = 0; c < N / 2; ++c) row[c] = ln.y[2 * c];
        ln.mu0[0] = ln.mu0[1] = ln.mu0[2] = 0.0;
    } else {
        P m[3];
        dipole<SYS, P>(sc, ln.y, m);
        ln.mu0[0] = m[0];
        ln.mu0[1] = m[1];
        ln.mu0[2] = m[2];
        if (jb.record) {
            row[0] = ln.mu0[0];
            row[jb.L] = ln.mu0[1];
            row[2 * (long long)jb.L] = ln.mu0[2];
        } else {
            row[0] = ln.mu0[0] * ln.mu0[0] + ln.mu0[1] * ln.mu0[1] + ln.mu0[2] * ln.mu0[2];
        }
    }
    ln.E0 = hamiltonian<SYS, P>(sc, ln.y);
    ln.t = 0.0;
    ln.k = 1;
    ln.steps = 0;
    ln.total_steps = 0;
    ln.reject = 0;
    ln.nrhs = 0;
    ln.prev = 0.0;
    ln.curr = 0.0;
    ln.active = 1;
    if (jb.L <= 1) {
        finish_item<N>(jb, ln, ITEM_OK, 0, 0.0f);
        return false;
    }
    {
        double f[N];
        rhs<SYS, P>(sc, ln.y, f);
#pragma unroll
        for (int i = 0; i < N; i += 2) ar.st2(0, i, f[i], f[i + 1]);
    }
    ln.nrhs = 1;
    ln.h = initial_step<SYS, P, N>(sc, ar, ln.y, jb.rtol, jb.atol, (jb.L - 1) * jb.dt, ln.nrhs);
    return true;
}

// out[j][i] = sum over the non-zeros q of tableau row R0 + j of A[R0 + j][q] * K_{col q}[i] for
// NR consecutive rows that share one column list (rows 16-18: B, E5, E3; rows 19-22: the dense
// output D rows), components i with i % STRIDE == 0 only (STRIDE 2: the coordinates).  The row
// is a compile-time constant, so columns, coefficients and arena offsets are immediates.
// Columns are loaded G at a time (a compiler fence after each group bounds the registers held)
// and each loaded column serves all NR rows; per row the sum runs over q in increasing order.
__device__ __forceinline__ void arena_fence();

template <int N, int R0, int NR, int G, int STRIDE>
__device__ __forceinline__ void rows_dot(const Arena<N>& ar, double (&out)[NR][N])
{
    constexpr int nnz = DOP853_SP_NNZ[R0];
#pragma unroll
    for (int j = 0; j < NR; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) out[j][i] = 0.0;
#pragma unroll
    for (int q0 = 0; q0 < nnz; q0 += G) {
        double v[G][N];
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (q0 + g < nnz) {
                if constexpr (STRIDE == 1) {
#pragma unroll
                    for (int i = 0; i < N; i += 2) ar.ld2(DOP853_SP_COL[R0][q0 + g], i, v[g][i], v[g][i + 1]);
                } else {
#pragma unroll
                    for (int i = 0; i < N; i += STRIDE) v[g][i] = ar.ld(DOP853_SP_COL[R0][q0 + g], i);
                }
            }
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (q0 + g < nnz) {
#pragma unroll
                for (int j = 0; j < NR; ++j)
#pragma unroll
                    for (int i = 0; i < N; i += STRIDE) out[j][i] += DOP853_SP_VAL[R0 + j][q0 + g] * v[g][i];
            }
        arena_fence();
    }
}

// Column groups of the tableau rows for a kernel resident at W wavefronts per SIMD: at two
// (256 registers) three columns at a time; at one, six (at most two memory round trips per
// row), which still fits the 256 architectural VGPRs without spilling into AGPRs and measured
// faster than whole rows.
template <int W>
struct Groups {
    static constexpr int stage = W == 1 ? 6 : 3, be = W == 1 ? 4 : 2, dense = W == 1 ? 6 : 3;
};

template <int N, int R, int G>
__device__ __forceinline__ void row_dot_c(const Arena<N>& ar, double* out)
{
    double o[1][N];
    rows_dot<N, R, 1, G, 1>(ar, o);
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = o[0][i];
}

// out = sum over the non-zero entries (m, a) of tableau row r (1..15, uniform across the
// wavefront) of a * K_m: one compile-time row per case
template <int N, int G>
__device__ __forceinline__ void row_dot(const Arena<N>& ar, int r, double* out)
{
    switch (r) {
    case 1: row_dot_c<N, 1, G>(ar, out); break;
    case 2: row_dot_c<N, 2, G>(ar, out); break;
    case 3: row_dot_c<N, 3, G>(ar, out); break;
    case 4: row_dot_c<N, 4, G>(ar, out); break;
    case 5: row_dot_c<N, 5, G>(ar, out); break;
    case 6: row_dot_c<N, 6, G>(ar, out); break;
    case 7: row_dot_c<N, 7, G>(ar, out); break;
    case 8: row_dot_c<N, 8, G>(ar, out); break;
    case 9: row_dot_c<N, 9, G>(ar, out); break;
    case 10: row_dot_c<N, 10, G>(ar, out); break;
    case 11: row_dot_c<N, 11, G>(ar, out); break;
    case 12: row_dot_c<N, 12, G>(ar, out); break;
    case 13: row_dot_c<N, 13, G>(ar, out); break;
    case 14: row_dot_c<N, 14, G>(ar, out); break;
    default: row_dot_c<N, 15, G>(ar, out); break;
    }
}

// Compiler-only barrier: stage vectors are re-read from the arena after it instead of being
// kept in registers across stages (which would rebuild the register-resident layout).
__device__ __forceinline__ void arena_fence() { __asm__ __volatile__("" ::: "memory"); }

// stages s0..s1 (inclusive): K_s = f(y + h sum_m A[s][m] K_m); noise: the largest FP32 noise
// bound of the stages' momentum derivatives
template <int SYS, typename P, int N, int W = 2>
__device__ __forceinline__ void stages(const SysConst& sc, const Arena<N>& ar, const double* y, double h, int s0,
                                       int s1, double& noise)
{
    noise = 0.0;
#pragma unroll 1
    for (int s = s0; s <= s1; ++s) {
        arena_fence();
        double yt[N], k[N];
        row_dot<N, Groups<W>::stage>(ar, s, yt);
#pragma unroll
        for (int i = 0; i < N; ++i) yt[i] = y[i] + h * yt[i];
        double nz;
        rhs<SYS, P>(sc, yt, k, nz);
        noise = fmax(noise, nz);
#pragma unroll
        for (int i = 0; i < N; i += 2) ar.st2(s, i, k[i], k[i + 1]);
    }
}

// Dense output of coordinate c (state index 2c) at x = (t_k - t) / h of an accepted step:
// DOP853's 7th-order continuous extension (Hairer's contd8) in Horner form.
__device__ __forceinline__ double dense_value(const double (&F)[7], double y0, double x)
{
    const double omx = 1.0 - x;
    double v = F[5] + x * F[6];
    v = F[4] + omx * v;
    v = F[3] + x * v;
    v = F[2] + omx * v;
    v = F[1] + x * v;
    v = F[0] + omx * v;
    return y0 + x * v;
}

template <int NQ>
__device__ __forceinline__ double dense_coord(const double (&Fc)[7][NQ], int c, double y0, double x)
{
    const double F[7] = {Fc[0][c], Fc[1][c], Fc[2][c], Fc[3][c], Fc[4][c], Fc[5][c], Fc[6][c]};
    return dense_value(F, y0, x);
}

// CF value mu(0) . mu(t_k) of a sample
template <typename P>
__device__ __forceinline__ double cf_value(const double* mu0, const P (&m)[3])
{
    return mu0[0] * (double)m[0] + mu0[1] * (double)m[1] + mu0[2] * (double)m[2];
}

// Checks of one sample in the reference's order (src/hawaii.c:2181-2236): a NaN dipole, a
// huge CF value, a jump by more than 1e10 from the previous sample (from sample 2 on).  A
// PR/mu series (record 2) has no check that ends it: the reference skips a sample whose dipole
// is NaN and integrates on (src/hawaii.c:3664-3677), leaving that sample of its zeroed series at
// 0 (store_sample).
template <typename P>
__device__ __forceinline__ int sample_status(const P (&m)[3], double curr, double prev, int k, int record)
{
    if (record == 2) return ITEM_OK;
    if (isnan((double)m[0]) || isnan((double)m[1]) || isnan((double)m[2])) return ITEM_FAIL_DIPOLE_NAN;
    if (fabs(curr) > 1e100) return ITEM_FAIL_CF_HUGE;
    if (k > 1 && fabs(curr / prev) > 1e10) return ITEM_FAIL_CF_JUMP;
    return ITEM_OK;
}

// sample k of an item whose row starts at `row`: mu(0) . mu(t_k), or the three dipole
// components (rows [3][L]) when the job records the series
template <typename P>
__device__ __forceinline__ void store_sample(double* row, int L, int record, int k, const P (&m)[3], double curr)
{
    if (record) {
        const bool skip = record == 2 && (isnan((double)m[0]) || isnan((double)m[1]) || isnan((double)m[2]));
        row[k] = skip ? 0.0 : (double)m[0];
        row[L + k] = skip ? 0.0 : (double)m[1];
        row[2 * (long long)L + k] = skip ? 0.0 : (double)m[2];
    } else {
        row[k] = curr;
    }
}

enum StepOutcome : int { STEP_ENDED = 0, STEP_REJECTED = 1, STEP_ACCEPTED = 2 };

// One attempted DOP853 step [t, t + h] of a lane: stages, error estimate, step-size control.
// STEP_ACCEPTED: yn holds the new state, h the step taken and fac11 the error factor; when a
// sample time falls in (t, t + h] (*dense) the dense-output coefficients Fc of the coordinates
// are computed as well.  STEP_ENDED: the item stopped (finish_item was called).
template <int SYS, typename P, int N, int W = 2>
__device__ __forceinline__ int dop853_attempt(const SysConst& sc, const JobDesc& jb, const Arena<N>& ar, Lane<N>& ln,
                                              double (&yn)[N], double& h_taken, double& fac11_out, bool& dense,
                                              double (&Fc)[7][N / 2])
{
    const double tend = (jb.L - 1) * jb.dt;
    dense = false;
    // CVode's mxstep test (src/trajectory.c:15, 62): max_steps accepted steps without reaching
    // the next sample time; kMaxTotalSteps bounds the step attempts of the item's whole life,
    // so every lane's loop ends.
    if (ln.steps >= jb.max_steps || ++ln.total_steps > kMaxTotalSteps) {
        finish_item<N>(jb, ln, ITEM_FAIL_STEPS, ln.k - 1, 0.0f);
        return STEP_ENDED;
    }
    if (0.1 * fabs(ln.h) <= fabs(ln.t) * kEPS) {
        finish_item<N>(jb, ln, ITEM_FAIL_STEPSIZE, ln.k - 1, 0.0f);
        return STEP_ENDED;
    }
    double h = ln.h;
    if (ln.t + 1.01 * h - tend > 0.0) h = tend - ln.t;

    // stages 1..11 and stage 12 = f(t + h, y_new) (tableau row 12 is the B row)
    double noise;
    stages<SYS, P, N, W>(sc, ar, ln.y, h, 1, 12, noise);
    ln.nrhs += 12;
    arena_fence();
    // y_new (B row) and the two error estimates (E5, E3 rows) share their columns
    static_assert(DOP853_ROW_E5 == DOP853_ROW_B + 1 && DOP853_ROW_E3 == DOP853_ROW_B + 2, "tableau row order");
    double err = 0.0, err2 = 0.0;
    {
        double be[3][N];
        rows_dot<N, DOP853_ROW_B, 3, Groups<W>::be, 1>(ar, be);
        // FP32 surfaces: the momentum tolerance is raised by the step's force rounding noise
        // times h, so the controller does not shrink the step to resolve FP32 noise (zero and
        // without effect in FP64)
        const double floor_p = fabs(h) * noise;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            yn[i] = ln.y[i] + h * be[0][i];
            const double sk = jb.atol + jb.rtol * fmax(fabs(ln.y[i]), fabs(yn[i])) + ((i & 1) ? floor_p : 0.0);
            if constexpr (sizeof(P) == 4) {
                // FP32 surfaces: one Newton-refined reciprocal per component (see rcp_nr)
                const double isk = rcp_nr(sk), e5 = be[1][i] * isk, e3 = be[2][i] * isk;
                err += e5 * e5;
                err2 += e3 * e3;
            } else {
                err += (be[1][i] / sk) * (be[1][i] / sk);
                err2 += (be[2][i] / sk) * (be[2][i] / sk);
            }
        }
    }
    double deno = err + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    err = fabs(h) * err * sqrt(1.0 / (N * deno));
    const double fac11 = sqrt(sqrt(sqrt(err)));
    if (!(err == err) || !(h == h)) {   // non-finite state: stop the item
        finish_item<N>(jb, ln, ITEM_FAIL_STEPSIZE, ln.k - 1, 0.0f);
        return STEP_ENDED;
    }
    if (err > 1.0) {
        ln.h = h / fmin(1.0 / 0.333, fac11 / 0.9);
        ln.reject = 1;
        return STEP_REJECTED;
    }
    ++ln.steps;
    h_taken = h;
    fac11_out = fac11;
    const double tnew = ln.t + h;
    if (ln.k < jb.L && ln.k * jb.dt <= tnew) {
        dense = true;
        double nz_dense;
        stages<SYS, P, N, W>(sc, ar, ln.y, h, 13, 15, nz_dense);
        ln.nrhs += 3;
        arena_fence();
        // F0 = dy, F1 = h K0 - dy, F2 = 2 dy - h (K12 + K0), F3..F6 = h (D K), for the
        // coordinates only (even state indices: the dipole and the R > Rcut test need no
        // momenta), kept in registers for all samples of the step
        constexpr int NQ = N / 2;
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            const int i = 2 * c;
            const double dy = yn[i] - ln.y[i];
            const double k0 = ar.ld(0, i);
            Fc[0][c] = dy;
            Fc[1][c] = h * k0 - dy;
            Fc[2][c] = 2.0 * dy - h * (ar.ld(12, i) + k0);
        }
        double f[4][N];
        rows_dot<N, DOP853_ROW_D, 4, Groups<W>::dense, 2>(ar, f);
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int c = 0; c < NQ; ++c) Fc[3 + d][c] = h * f[d][2 * c];
    }
    return STEP_ACCEPTED;
}

// After an accepted step that did not end the item: y <- y_new, f(y) <- K12
// (first-same-as-last), t <- t + h and the next step size.
template <int N>
__device__ __forceinline__ void advance(const JobDesc& jb, const Arena<N>& ar, Lane<N>& ln, const double (&yn)[N],
                                        double h, double fac11)
{
    const double tend = (jb.L - 1) * jb.dt;
    arena_fence();
#pragma unroll
    for (int i = 0; i < N; i += 2) {
        ln.y[i] = yn[i];
        ln.y[i + 1] = yn[i + 1];
        double a, b;
        ar.ld2(12, i, a, b);
        ar.st2(0, i, a, b);
    }
    ln.t = ln.t + h;
    const double fac = fmax(1.0 / 6.0, fmin(1.0 / 0.333, fac11 / 0.9));
    double hnew = h / fac;
    if (fabs(hnew) > tend) hnew = tend;
    if (ln.reject) {
        hnew = fmin(fabs(hnew), fabs(h));
        ln.reject = 0;
    }
    ln.h = hnew;
}

// One attempted DOP853 step for a lane that records its own samples (the plain engine: one
// item per lane for its whole life).  Returns true when the lane's item has ended.
template <int SYS, typename P, int N>
__device__ __forceinline__ bool dop853_step(const SysConst& sc, const JobDesc* __restrict__ jobs, const Arena<N>& ar,
                                            Lane<N>& ln)
{
    constexpr int NQ = N / 2;
    const JobDesc& jb = jobs[ln.job];
    double yn[N], Fc[7][NQ], h, fac11;
    bool dense;
    const int outcome = dop853_attempt<SYS, P, N>(sc, jb, ar, ln, yn, h, fac11, dense, Fc);
    if (outcome == STEP_ENDED) return true;
    if (outcome == STEP_REJECTED) return false;
    // accepted step [t, t + h]: record every sample t_k in (t, t + h]
    const double tnew = ln.t + h;
    bool ended = false;
    int status = ITEM_OK;
    float drift = 0.0f;
    if (dense) {
        double* row = jb.rows + ln.local * (long long)jb.L * row_width(jb.record);
        while (ln.k < jb.L && ln.k * jb.dt <= tnew) {
            const double tk = ln.k * jb.dt;
            double ys[N];
            if (tk == tnew) {
#pragma unroll
                for (int i = 0; i < N; ++i) ys[i] = yn[i];
            } else {
                const double x = (tk - ln.t) / h;
#pragma unroll
                for (int i = 1; i < N; i += 2) ys[i] = 0.0;
#pragma unroll
                for (int c = 0; c < NQ; ++c) ys[2 * c] = dense_coord<NQ>(Fc, c, ln.y[2 * c], x);
            }
            P m[3];
            dipole<SYS, P>(sc, ys, m);
            const double curr = cf_value<P>(ln.mu0, m);
            status = sample_status<P>(m, curr, ln.curr, ln.k, jb.record);
            ln.prev = ln.curr;
            ln.curr = curr;
            if (status != ITEM_OK) {
                ended = true;
                break;
            }
            store_sample<P>(row, jb.L, jb.record, ln.k, m, curr);
            ln.steps = 0;
            if (ys[4] > jb.Rcut || ln.k + 1 >= jb.L) {
                // energy check of make_step (src/trajectory.c:167-180), at the end of the step
                // that holds the last sample
                drift = (float)fabs(hamiltonian<SYS, P>(sc, yn) - ln.E0);
                ended = true;
                break;
            }
            ++ln.k;
        }
    }
    if (ended) {
        // on success sample ln.k was stored; a failed check stops before storing it
        finish_item<N>(jb, ln, status, status == ITEM_OK ? ln.k : ln.k - 1, drift);
        return true;
    }
    advance<N>(jb, ar, ln, yn, h, fac11);
    return false;
}

// ------------------------------------------------------------------ wavefront-shared sampling
// In the batched engine the samples of a step are evaluated by the whole wavefront.  A lane's
// step covers a lane-dependent number of sample times (none inside a collision, dozens in free
// flight), so a lane-serial loop would run as many dipole passes as the wavefront's busiest
// lane while the others idle.  Instead every lane with samples publishes its step (state,
// dense-output coefficients, CF context) in arena slots 1.. (free once the step's stages are
// combined) and the wavefront evaluates all pending samples 64 at a time: sample g of the
// wavefront's list belongs to the lane whose exclusive prefix count is the largest <= g.
// Values, checks and the stored samples are those of the lane-serial loop: sample k compares
// with sample k - 1 (the previous pass's last lane, or the lane's CF value at the step start)
// and only samples before the first failing one are stored.
struct WaveShare {
    int pre[65];      // exclusive prefix of the lanes' sample counts; pre[64] = total
    int fail[64];     // (index in the step << 3) | status of the lane's first failing sample
    double last[64];  // CF value of the lane's last sample
};

// Fields of a lane's step record (field f at arena slot 1 + f / N, component f % N), in pairs
// that one 16-byte access moves: the sample-time scalars [0, CF0), the CF scalars [CF0, YN),
// the coordinates at the step's end (used only for a sample at exactly t + h), then per
// coordinate c its value at the step's start followed by its seven dense-output coefficients
// (Y + 8c .. Y + 8c + 7).
template <int NQ>
struct StepRecord {
    static constexpr int T = 0, H = 1, TNEW = 2, DT = 3, K0 = 4, CF0 = 6, MU0 = 6, CURR = 9, ROW = 10, L = 11,
                         RECORD = 12, YN = 14, Y = YN + (NQ + 1) / 2 * 2, SIZE = Y + 8 * NQ;
};

// fields [f0, f0 + n) of the record of lane `owner` (f0 and n even)
template <int N, int NREC>
__device__ __forceinline__ void rec_load_block(const Arena<N>& ar, int owner, int f0, int n, double (&r)[NREC])
{
#pragma unroll
    for (int f = f0; f < f0 + n; f += 2) ar.ld2_lane(owner, 1 + f / N, f % N, r[f], r[f + 1]);
}

// Memory written by one lane (LDS, the wavefront's arena block) is visible to the others
// after this; every lane of the wavefront executes it.  A wavefront's own memory operations
// are performed in order and its lanes share one vector L1, so the wavefront-scope fence only
// keeps the compiler from moving accesses across it (it emits no wait).
__device__ __forceinline__ void wave_share_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Evaluate, check and store the n pending samples of every lane (n = 0: none).  Called by all
// 64 lanes of the wavefront with the same control flow.
template <int SYS, typename P, int N>
__device__ __forceinline__ void shared_samples(const SysConst& sc, const Arena<N>& ar, WaveShare& ws, int n)
{
    constexpr int NQ = N / 2;
    using R = StepRecord<NQ>;
    const int lane = (int)__lane_id();
    int incl = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    const int total = __shfl(incl, 63, 64);
    if (total == 0) return;
    ws.pre[lane] = incl - n;
    if (lane == 0) ws.pre[64] = total;
    ws.fail[lane] = 0x7fffffff;
    wave_share_sync();
    double carry = 0.0;   // CF value of the previous pass's last sample
#pragma unroll 1
    for (int base = 0; base < total; base += 64) {
        const int g = base + lane;
        const bool valid = g < total;
        int owner = 0, q = 0, k = 0, cnt = 0, rowL = 0, record = 0;
        double curr = 0.0, prev_start = 0.0;
        double* row = nullptr;
        P m[3] = {P(0), P(0), P(0)};
        if (valid) {
            int lo = 0, hi = 63;
#pragma unroll
            for (int it = 0; it < 6; ++it) {
                const int mid = (lo + hi + 1) >> 1;
                if (ws.pre[mid] <= g)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            owner = lo;
            q = g - ws.pre[owner];
            cnt = ws.pre[owner + 1] - ws.pre[owner];
            double r[R::SIZE];
            rec_load_block<N>(ar, owner, 0, R::CF0, r);
            k = (int)r[R::K0] + q;
            const double tk = k * r[R::DT];
            double ys[N];
#pragma unroll
            for (int i = 1; i < N; i += 2) ys[i] = 0.0;
            if (tk == r[R::TNEW]) {
                rec_load_block<N>(ar, owner, R::YN, R::Y - R::YN, r);
#pragma unroll
                for (int c = 0; c < NQ; ++c) ys[2 * c] = r[R::YN + c];
            } else {
                const double x = sizeof(P) == 4 ? (tk - r[R::T]) * rcp_nr(r[R::H]) : (tk - r[R::T]) / r[R::H];
#pragma unroll
                for (int c = 0; c < NQ; ++c) {
                    rec_load_block<N>(ar, owner, R::Y + 8 * c, 8, r);
                    const double F[7] = {r[R::Y + 8 * c + 1], r[R::Y + 8 * c + 2], r[R::Y + 8 * c + 3],
                                         r[R::Y + 8 * c + 4], r[R::Y + 8 * c + 5], r[R::Y + 8 * c + 6],
                                         r[R::Y + 8 * c + 7]};
                    ys[2 * c] = dense_value(F, r[R::Y + 8 * c], x);
                }
            }
            rec_load_block<N>(ar, owner, R::CF0, R::YN - R::CF0, r);
            row = __builtin_bit_cast(double*, r[R::ROW]);
            rowL = (int)r[R::L];
            record = (int)r[R::RECORD];
            if constexpr (SYS == SYS_CH4CO2) {
                // record 3: the sample's coordinates; its dipole and CF value come later
#pragma unroll
                for (int c = 0; c < NQ; ++c) row[(long long)k * NQ + c] = ys[2 * c];
            } else {
                dipole<SYS, P>(sc, ys, m);
                const double mu0[3] = {r[R::MU0], r[R::MU0 + 1], r[R::MU0 + 2]};
                curr = cf_value<P>(mu0, m);
            }
            prev_start = r[R::CURR];
        }
        const double up = __shfl_up(curr, 1, 64);
        const double prev = q == 0 ? prev_start : (lane > 0 ? up : carry);
        carry = __shfl(curr, 63, 64);
        if (valid && SYS != SYS_CH4CO2) {
            const int st = sample_status<P>(m, curr, prev, k, record);
            if (st != ITEM_OK) atomicMin(&ws.fail[owner], (q << 3) | st);
        }
        wave_share_sync();
        if (valid) {
            if (SYS != SYS_CH4CO2 && q < (ws.fail[owner] >> 3)) store_sample<P>(row, rowL, record, k, m, curr);
            if (q == cnt - 1) ws.last[owner] = curr;
        }
    }
    wave_share_sync();
}

// One attempted DOP853 step of every lane of a wavefront (the batched engine); lanes without
// an item take part in the sample evaluation only.
template <int SYS, typename P, int N, int W>
__device__ __forceinline__ void dop853_step_shared(const SysConst& sc, const JobDesc* __restrict__ ring,
                                                   const Arena<N>& ar, Lane<N>& ln, WaveShare& ws)
{
    constexpr int NQ = N / 2;
    using R = StepRecord<NQ>;
    static_assert(1 + (R::SIZE - 1) / N < 12, "the step record must leave K0 and K12 alone");
    double yn[N], h = 0.0, fac11 = 0.0;
    int outcome = STEP_ENDED, n = 0;
    bool end = false;
    if (ln.active) {
        const JobDesc& jb = ring[ln.job];
        double Fc[7][NQ];
        bool dense;
        outcome = dop853_attempt<SYS, P, N, W>(sc, jb, ar, ln, yn, h, fac11, dense, Fc);
        if (outcome == STEP_ACCEPTED && dense) {
            // the samples of this step, up to the one that ends the item (R > Rcut on the
            // dense output, or the last sample index)
            const double tnew = ln.t + h;
            // On [t, tnew] the dense output of R stays within sum_d |Fc[d][2]| of R(t) (x and
            // 1 - x lie in [0, 1]); when that bound, with slack for rounding, and R(tnew) are
            // below Rcut no sample can end the item and the count follows from the sample times.
            double span = 0.0;
#pragma unroll
            for (int d = 0; d < 7; ++d) span += fabs(Fc[d][2]);
            const double rmax = ln.y[4] + span * (1.0 + 1e-12) + 1e-12 * fabs(ln.y[4]);
            if (rmax < jb.Rcut && yn[4] < jb.Rcut) {
                // last sample index kf <= L - 1 with kf dt <= tnew (the loop's own predicate)
                int kf = (int)fmin(floor(tnew / jb.dt), (double)(jb.L - 1));
                while (kf + 1 <= jb.L - 1 && (kf + 1) * jb.dt <= tnew) ++kf;
                while (kf > ln.k && kf * jb.dt > tnew) --kf;
                n = kf - ln.k + 1;
                end = kf + 1 >= jb.L;
            } else {
                for (int k = ln.k; k < jb.L && k * jb.dt <= tnew; ++k) {
                    ++n;
                    const double tk = k * jb.dt;
                    const double Rk = tk == tnew ? yn[4] : dense_coord<NQ>(Fc, 2, ln.y[4], (tk - ln.t) / h);
                    if (Rk > jb.Rcut || k + 1 >= jb.L) {
                        end = true;
                        break;
                    }
                }
            }
            double r[R::SIZE];
            r[R::T] = ln.t;
            r[R::H] = h;
            r[R::TNEW] = tnew;
            r[R::DT] = jb.dt;
            r[R::MU0] = ln.mu0[0];
            r[R::MU0 + 1] = ln.mu0[1];
            r[R::MU0 + 2] = ln.mu0[2];
            r[R::CURR] = ln.curr;
            r[R::K0] = (double)ln.k;
            r[R::ROW] = __builtin_bit_cast(double, jb.rows + ln.local * (long long)jb.L * row_width(jb.record));
            r[R::L] = (double)jb.L;
            r[R::RECORD] = (double)jb.record;
            r[R::K0 + 1] = 0.0;
            r[R::RECORD + 1] = 0.0;
#pragma unroll
            for (int f = R::YN + NQ; f < R::Y; ++f) r[f] = 0.0;
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                r[R::YN + c] = yn[2 * c];
                r[R::Y + 8 * c] = ln.y[2 * c];
#pragma unroll
                for (int d = 0; d < 7; ++d) r[R::Y + 8 * c + 1 + d] = Fc[d][c];
            }
#pragma unroll
            for (int f = 0; f < R::SIZE; f += 2) ar.st2(1 + f / N, f % N, r[f], r[f + 1]);
        }
    }
    shared_samples<SYS, P, N>(sc, ar, ws, n);
    if (outcome != STEP_ACCEPTED) return;
    const JobDesc& jb = ring[ln.job];
    if (n > 0) {
        const int lane = (int)__lane_id();
        const int f = ws.fail[lane];
        if (f != 0x7fffffff) {
            // a failed check stops before storing its sample
            finish_item<N>(jb, ln, f & 7, ln.k + (f >> 3) - 1, 0.0f);
            return;
        }
        ln.steps = 0;
        if (end) {
            // energy check of make_step (src/trajectory.c:167-180), at the end of the step
            // that holds the last sample
            const float drift = (float)fabs(hamiltonian<SYS, P>(sc, yn) - ln.E0);
            finish_item<N>(jb, ln, ITEM_OK, ln.k + n - 1, drift);
            return;
        }
        ln.k += n;
        ln.prev = ln.curr;
        ln.curr = ws.last[lane];
    }
    // Mixed precision: an outer radial turning point (pR from > 0 to <= 0) means the pair is
    // bound in an orbiting complex, whose trajectory depends on the last bit of the arithmetic;
    // the item is handed over for integration in FP64 from its initial condition.
    if (jb.orbit_items && ln.y[5] > 0.0 && yn[5] <= 0.0) {
        const int at = atomicAdd(jb.orbit_count, 1);
        jb.orbit_items[at] = (int)ln.local;
        finish_item<N>(jb, ln, ITEM_ORBIT, ln.k - 1, 0.0f);
        return;
    }
    advance<N>(jb, ar, ln, yn, h, fac11);
}

// ------------------------------------------------------------------ kernels
// CO2-Ar surface tables are staged into LDS once per workgroup: every right-hand side reads
// 6 Legendre slots x 8 coefficients at a lane-dependent radial branch.
template <int SYS, typename P>
__device__ __forceinline__ SysConst stage_tables(const SysConst& sc, Co2ArTables<P>* s_tab)
{
    SysConst out = sc;
    if constexpr (SYS == SYS_CO2AR) {
        const double* src = (const double*)sc.tables;
        double* dst = (double*)s_tab;
        for (unsigned i = threadIdx.x; i < sizeof(Co2ArTables<P>) / sizeof(double); i += blockDim.x) dst[i] = src[i];
        __syncthreads();
        out.tables = s_tab;
    }
    return out;
}

// plain: lane g owns item g; ws holds one arena block per wavefront of the grid
template <int SYS>
__global__ void __launch_bounds__(256, HH_MIN_WAVES) plain_cf_kernel(SysConst sc, const JobDesc* __restrict__ jobs, int njobs,
                                                       long long total_items, double* __restrict__ ws)
{
    constexpr int N = SysTraits<SYS>::QP;
    __shared__ Co2ArTables<double> s_tab;
    const SysConst scl = stage_tables<SYS, double>(sc, &s_tab);
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const Arena<N> ar = wave_arena<N>(ws, (long long)blockIdx.x * 4 + (threadIdx.x >> 6));
    if (g >= total_items) return;
    Lane<N> ln;
    const int j = find_job(jobs, njobs, g);
    if (!init_item<SYS, double, N>(scl, jobs[j], j, g - jobs[j].item_begin, ar, ln)) return;
    while (!dop853_step<SYS, double, N>(scl, jobs, ar, ln)) {
    }
}

// ------------------------------------------------------------------ resident lane pool
template <int N>
__device__ __forceinline__ void restore_lane(const LaneSlots& sl, long long s, const Arena<N>& ar, Lane<N>& ln)
{
    const long long S = sl.S;
    ln.job = sl.i[s];
    ln.active = ln.job >= 0;
    if (!ln.active) return;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        ln.y[i] = sl.d[i * S + s];
        ar.st(0, i, sl.d[(N + i) * S + s]);
    }
    const double* d = sl.d + 2 * N * S + s;
    ln.t = d[0];
    ln.h = d[S];
    ln.E0 = d[2 * S];
    ln.mu0[0] = d[3 * S];
    ln.mu0[1] = d[4 * S];
    ln.mu0[2] = d[5 * S];
    ln.prev = d[6 * S];
    ln.curr = d[7 * S];
    ln.k = sl.i[S + s];
    ln.steps = sl.i[2 * S + s];
    ln.total_steps = sl.i[3 * S + s];
    ln.nrhs = sl.i[4 * S + s];
    ln.reject = sl.i[5 * S + s];
    ln.local = sl.local[s];
}

template <int N>
__device__ __forceinline__ void save_lane(const LaneSlots& sl, long long s, const Arena<N>& ar, const Lane<N>& ln)
{
    const long long S = sl.S;
    if (!ln.active) {
        sl.i[s] = -1;
        return;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        sl.d[i * S + s] = ln.y[i];
        sl.d[(N + i) * S + s] = ar.ld(0, i);
    }
    double* d = sl.d + 2 * N * S + s;
    d[0] = ln.t;
    d[S] = ln.h;
    d[2 * S] = ln.E0;
    d[3 * S] = ln.mu0[0];
    d[4 * S] = ln.mu0[1];
    d[5 * S] = ln.mu0[2];
    d[6 * S] = ln.prev;
    d[7 * S] = ln.curr;
    sl.i[s] = ln.job;
    sl.i[S + s] = ln.k;
    sl.i[2 * S + s] = ln.steps;
    sl.i[3 * S + s] = ln.total_steps;
    sl.i[4 * S + s] = ln.nrhs;
    sl.i[5 * S + s] = ln.reject;
    sl.local[s] = ln.local;
}

// ring slot of the job holding global item g: jobs [first_seq, last_seq] are in flight, their
// item ranges are consecutive and increasing with the sequence number
__device__ __forceinline__ int ring_job(const JobDesc* __restrict__ ring, const PoolCtl& c, unsigned long long g)
{
    long long lo = c.first_seq, hi = c.last_seq;
    while (lo < hi) {
        const long long mid = (lo + hi + 1) >> 1;
        if ((long long)g >= ring[mid % c.ring].item_begin)
            lo = mid;
        else
            hi = mid - 1;
    }
    return (int)(lo % c.ring);
}

// One slice of the batched engine.  Every lane slot restores its item (if any), integrates
// and refills from the queue until the slice's time is up or no work is left, and saves its
// state at a step boundary, so slices can be separated by admission of new jobs, reductions
// of finished ones and repacking of the slots.  The loop is wavefront-uniform (lanes without
// an item stay in it to evaluate samples of the others) and ends: the deadline and the queue
// bound it, and every item ends within kMaxTotalSteps attempts.
template <int SYS, typename P, int BS, int W>
__global__ void __launch_bounds__(BS, W) pool_cf_kernel(SysConst sc, const JobDesc* __restrict__ ring,
                                                                   PoolCtl ctl, LaneSlots sl, double* __restrict__ ws)
{
    constexpr int N = SysTraits<SYS>::QP;
    __shared__ Co2ArTables<P> s_tab;
    __shared__ WaveShare s_share[BS / 64];
    const SysConst scl = stage_tables<SYS, P>(sc, &s_tab);
    const long long wave = (long long)blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
    const Arena<N> ar = wave_arena<N>(ws, wave);
    WaveShare& share = s_share[threadIdx.x >> 6];
    const long long slot = (long long)blockIdx.x * BS + threadIdx.x;
    const long long t_end = wall_clock64() + ctl.ticks;
    Lane<N> ln;
    restore_lane<N>(sl, slot, ar, ln);
    bool queue_open = true;   // the queue only grows between slices
    for (;;) {
        if (wall_clock64() >= t_end) break;
        if (!ln.active && queue_open) {
            bool got = false;
            while (!got) {
                if (__hip_atomic_load(ctl.next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= ctl.total) break;
                const unsigned long long g = atomicAdd(ctl.next, 1ull);
                if (g >= ctl.total) break;
                int j = ring_job(ring, ctl, g);
                unsigned long long item = g;
                if (ring[j].order) {
                    item = (unsigned long long)(ring[j].order_base +
                                                ring[j].order[(long long)g - ring[j].order_base]);
                    j = ring_job(ring, ctl, item);
                }
                got = init_item<SYS, P, N>(scl, ring[j], j, (long long)item - ring[j].item_begin, ar, ln);
            }
            queue_open = got;
        }
        if (__ballot(ln.active) == 0) break;
        dop853_step_shared<SYS, P, N, W>(scl, ring, ar, ln, share);
    }
    save_lane<N>(sl, slot, ar, ln);
}

// Scheduling keys of every item of a job.
//   what = ORDER_TRANSIT_KEY: the straight-line time (in samples, at most L - 1) from the initial
//     point to the Rcut sphere along the item's initial velocity, forward for the first half and
//     momentum-inverted for the second.  Trajectory cost grows with its duration, so handing out
//     items in decreasing key order starts the long ones first.
//   what = ORDER_IMPACT_KEY: the squared impact parameter b^2 = L^2 / p^2 of the relative motion.
//   what = ORDER_TIERED_KEY: the transit time plus L for the items whose entry of `keys` is 1 on
//     input (the front tier, marked by order_front_kernel), so they are handed out first.
// Only the order of work changes; every item's result is the same.

__global__ void __launch_bounds__(256) order_keys_kernel(const double* __restrict__ ic, long long ntraj, int halves,
                                                         double mu, double Rcut, double dt, int L, int what,
                                                         float* __restrict__ keys, int* __restrict__ idx,
                                                         int idx_offset)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= ntraj * halves) return;
    const long long t = i / halves;
    const double pPhi = ic[1 * ntraj + t], Theta = ic[2 * ntraj + t], pTheta = ic[3 * ntraj + t];
    const double R = ic[4 * ntraj + t], pR = ic[5 * ntraj + t];
    const double sT = sin(Theta);
    const double L2 = pTheta * pTheta + (sT != 0.0 ? pPhi * pPhi / (sT * sT) : 0.0);
    const double p2 = pR * pR + L2 / (R * R);
    const double p = sqrt(p2);
    double key = (double)(L - 1);
    if (what == ORDER_IMPACT_KEY) {
        key = p2 > 0.0 ? L2 / p2 : 0.0;
    } else if (p > 0.0) {
        const double rr = ((i % halves) == 1 ? -1.0 : 1.0) * R * pR / p;   // r . u
        const double disc = rr * rr - R * R + Rcut * Rcut;
        const double tx = disc > 0.0 ? (-rr + sqrt(disc)) * mu / p : 0.0;
        key = fmin(tx / dt, (double)(L - 1));
    }
    if (what == ORDER_TIERED_KEY && keys[i] != 0.0f) key += (double)L;
    keys[i] = (float)key;
    idx[i] = idx_offset + (int)i;
}

// keys[order[r]] = 1 for the first n_front ranks of `order` (a permutation of [0, n)), 0 for
// the rest.
__global__ void __launch_bounds__(256) order_front_kernel(const int* __restrict__ order, long long n, long long n_front,
                                                          float* __restrict__ keys)
{
    const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    keys[order[r]] = r < n_front ? 1.0f : 0.0f;
}

// Batched surface evaluation at phase points y (SoA [QP][n], FP64): what=0 the potential
// (pes_lab, Hartree), 1 the dipole (dipole_lab, 3 x n), 2 the Hamiltonian, 3 the right-hand side
// of Hamilton's equations (QP x n).
template <int SYS>
__global__ void __launch_bounds__(256) eval_kernel(SysConst sc, int what, const double* __restrict__ y, long long n,
                                                   double* __restrict__ out)
{
    constexpr int N = SysTraits<SYS>::QP;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double v[N];
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = y[k * n + i];
    if (what == 0) {
        out[i] = pes_value<SYS, double>(sc, v);
    } else if (what == 1) {
        double m[3];
        dipole<SYS, double>(sc, v, m);
        out[i] = m[0];
        out[n + i] = m[1];
        out[2 * n + i] = m[2];
    } else if (what == 2) {
        out[i] = hamiltonian<SYS, double>(sc, v);
    } else if (what == 4) {
        if constexpr (SYS != SYS_CH4CO2) out[i] = dipole_rate_sq<SYS>(sc, v);
    } else {
        double f[N];
        rhs<SYS, double>(sc, v, f);
#pragma unroll
        for (int k = 0; k < N; ++k) out[k * n + i] = f[k];
    }
}

// Repack the active lane slots of `src` into `dst`: active slot number r (in slot order) goes
// to wavefront r % W, lane r / W, so the remaining items spread over all W wavefronts of the
// pool (one per wavefront once fewer than W are left) instead of sharing a few diverging ones.
// Pass 1 counts active slots per 256-slot block; pass 2 scans the counts (one workgroup);
// pass 3 moves every active slot and marks the others empty.
__global__ void __launch_bounds__(256) repack_count_kernel(const int* __restrict__ job, long long S, int* __restrict__ counts)
{
    const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
    const int a = (s < S && job[s] >= 0) ? 1 : 0;
    const unsigned long long m = __ballot(a);
    __shared__ int wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ void __launch_bounds__(1024) repack_scan_kernel(int* __restrict__ counts, int nblocks, int* __restrict__ total)
{
    __shared__ int part[1024];
    const int per = (nblocks + 1023) / 1024;
    const int b0 = threadIdx.x * per;
    int acc = 0;
    for (int b = b0; b < b0 + per && b < nblocks; ++b) acc += counts[b];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - acc;   // exclusive prefix of this thread's range
    for (int b = b0; b < b0 + per && b < nblocks; ++b) {
        const int c = counts[b];
        counts[b] = run;
        run += c;
    }
    if (threadIdx.x == 1023) *total = part[1023];
}

template <int N>
__global__ void __launch_bounds__(256) repack_move_kernel(LaneSlots src, LaneSlots dst, const int* __restrict__ offsets,
                                                          long long W)
{
    const long long S = src.S;
    const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
    const int a = (s < S && src.i[s] >= 0) ? 1 : 0;
    // rank inside the block: wave ballot prefix + preceding waves of the block
    const unsigned long long m = __ballot(a);
    const int lane = threadIdx.x & 63;
    __shared__ int wsum[4];
    if (lane == 0) wsum[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (!a) return;   // dst was cleared beforehand (slots_clear_kernel)
    int rank = offsets[blockIdx.x] + __popcll(m & ((1ull << lane) - 1ull));
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) rank += wsum[w];
    const long long d = (long long)(rank % W) * 64 + rank / W;
    for (int f = 0; f < 2 * N + 8; ++f) dst.d[f * S + d] = src.d[f * S + s];
    for (int f = 0; f < 6; ++f) dst.i[f * S + d] = src.i[f * S + s];
    dst.local[d] = src.local[s];
}

// right-hand sides spent so far on one job: its items still in lanes (slot counters) plus its
// finished items (per-item counters, zeroed at admission); one partial per block
__global__ void __launch_bounds__(256) pool_work_kernel(LaneSlots sl, int ring_slot, const int* __restrict__ item_nrhs,
                                                         long long n_items, unsigned long long* __restrict__ out)
{
    __shared__ unsigned long long part[4];
    const long long S = sl.S;
    unsigned long long w = 0;
    for (long long s = (long long)blockIdx.x * 256 + threadIdx.x; s < S; s += (long long)gridDim.x * 256)
        if (sl.i[s] == ring_slot) w += (unsigned long long)sl.i[4 * S + s];
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n_items; t += (long long)gridDim.x * 256)
        w += (unsigned long long)item_nrhs[t];
    for (int off = 32; off > 0; off >>= 1) w += __shfl_down(w, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, part[0] + part[1] + part[2] + part[3]);
}

hipError_t run_pool_work(const LaneSlots& sl, int ring_slot, const int* item_nrhs, long long n_items,
                         unsigned long long* out, hipStream_t st)
{
    hipLaunchKernelGGL(pool_work_kernel, dim3(256), dim3(256), 0, st, sl, ring_slot, item_nrhs, n_items, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) slots_clear_kernel(int* __restrict__ job, long long S)
{
    const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
    if (s < S) job[s] = -1;
}

// ------------------------------------------------------------------ ensemble reduction
__device__ __forceinline__ void neumaier_add(double& s, double& c, double v)
{
    const double t = s + v;
    if (fabs(s) >= fabs(v))
        c += (s - t) + v;
    else
        c += (v - t) + s;
    s = t;
}

// partial[c][k] = sum over trajectories of chunk c of ppf * crln[k]; column-parallel, rows
// read coalesced along k; entries past an item's length are zero and never read.  The items are
// taken kReduceUnroll at a time: their lengths and statuses (wavefront-uniform) and row values
// are all loaded first and then added in item order, so a wavefront keeps kReduceUnroll row
// loads in flight.  A row is read through a buffer resource whose size is the item's length
// (zero for an item that does not count or lies past the chunk): the load of a column past it
// returns 0 without a branch.  Such a column adds +0.0, which leaves the Neumaier sum and its
// compensation as they are, so the sum is that of the item-serial loop.
constexpr int kReduceUnroll = 16;

__device__ __forceinline__ double row_value(const double* row, int n, int k)
{
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(row), 0, n * 8, kBufferConfig);
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, k * 8, 0, 0));
}

template <int HALVES>
__device__ __forceinline__ double reduce_value(const JobDesc& jb, long long t, long long t1, int k)
{
    const bool in = t < t1;
    const long long tc = in ? t : t1 - 1;   // a valid item for the (discarded) status and length reads
    if constexpr (HALVES == 2) {
        const long long a = 2 * tc, b = 2 * tc + 1;
        const bool use = in && !dropped(jb.status[a], jb.status[b]);
        const double cf = row_value(jb.rows + a * jb.L, use ? jb.len[a] : 0, k);
        const double cb = row_value(jb.rows + b * jb.L, use ? jb.len[b] : 0, k);
        return 0.5 * (cf + cb) * kALU * kALU * kALU;
    } else {
        const bool use = in && !dropped(jb.status[tc], jb.status[tc]);
        const double cf = row_value(jb.rows + tc * jb.L, use ? jb.len[tc] : 0, k);
        return cf * kALU * kALU * kALU;
    }
}

template <int HALVES>
__device__ __forceinline__ void reduce_chunk(const JobDesc& jb, long long t0, long long t1, int k, double ppf,
                                             double& s, double& cc)
{
    for (long long t = t0; t < t1; t += kReduceUnroll) {
        double v[kReduceUnroll];
#pragma unroll
        for (int u = 0; u < kReduceUnroll; ++u) v[u] = reduce_value<HALVES>(jb, t + u, t1, k);
#pragma unroll
        for (int u = 0; u < kReduceUnroll; ++u) neumaier_add(s, cc, ppf * v[u]);
    }
}

__global__ void __launch_bounds__(256) cf_reduce_partial_kernel(JobDesc jb, long long chunk, double ppf,
                                                                double* __restrict__ psum,
                                                                double* __restrict__ pcomp)
{
    __shared__ int s_maxlen;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const long long t0 = (long long)c * chunk;
    long long t1 = t0 + chunk;
    if (t1 > jb.ntraj) t1 = jb.ntraj;
    // the longest row of the chunk: a workgroup whose columns all lie past it only writes zeros
    // (what adding +0.0 for every item gives)
    if (threadIdx.x == 0) s_maxlen = 0;
    __syncthreads();
    {
        int m = 0;
        for (long long t = t0 + threadIdx.x; t < t1; t += blockDim.x)
            m = max(m, jb.halves == 2 ? max(jb.len[2 * t], jb.len[2 * t + 1]) : jb.len[t]);
        atomicMax(&s_maxlen, m);
    }
    __syncthreads();
    if (k >= jb.L) return;
    double s = 0.0, cc = 0.0;
    if ((int)(blockIdx.x * blockDim.x) < s_maxlen) {
        if (jb.halves == 2)
            reduce_chunk<2>(jb, t0, t1, k, ppf, s, cc);
        else
            reduce_chunk<1>(jb, t0, t1, k, ppf, s, cc);
    }
    psum[(long long)c * jb.L + k] = s;
    pcomp[(long long)c * jb.L + k] = cc;
}

__global__ void __launch_bounds__(256) cf_reduce_final_kernel(int L, int nchunks, long long ntraj,
                                                              const double* __restrict__ psum,
                                                              const double* __restrict__ pcomp,
                                                              double* __restrict__ cf_out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L) return;
    double s = 0.0, cc = 0.0;
    // kReduceUnroll chunks' partials loaded at once, added in chunk order (a chunk index past the
    // end reads the last chunk again and adds +0.0)
    for (int c = 0; c < nchunks; c += kReduceUnroll) {
        double ps[kReduceUnroll], pc[kReduceUnroll];
#pragma unroll
        for (int u = 0; u < kReduceUnroll; ++u) {
            const bool in = c + u < nchunks;
            const long long o = (long long)(in ? c + u : nchunks - 1) * L + k;
            ps[u] = in ? psum[o] : 0.0;
            pc[u] = in ? pcomp[o] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kReduceUnroll; ++u) {
            neumaier_add(s, cc, ps[u]);
            cc += pc[u];
        }
    }
    cf_out[k] = (s + cc) / (double)ntraj;
}

// CORRELATION_ARRAY: the same sums with a weight per trajectory and satellite temperature,
// wr[t * ntemp + st] (Boltzmann reweight times the temperature's partition-function ratio);
// blockIdx.z is the temperature.  Partial sums psum[(st * nchunks + c) * L + k].
__global__ void __launch_bounds__(256) cf_reduce_weighted_partial_kernel(JobDesc jb, long long chunk, int ntemp,
                                                                         const double* __restrict__ wr,
                                                                         double* __restrict__ psum,
                                                                         double* __restrict__ pcomp)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y, st = blockIdx.z;
    if (k >= jb.L) return;
    const long long t0 = (long long)c * chunk;
    long long t1 = t0 + chunk;
    if (t1 > jb.ntraj) t1 = jb.ntraj;
    double s = 0.0, cc = 0.0;
    for (long long t = t0; t < t1; ++t) {
        double v;
        if (jb.halves == 2) {
            const long long a = 2 * t, b = 2 * t + 1;
            if (dropped(jb.status[a], jb.status[b])) continue;
            const double cf = k < jb.len[a] ? jb.rows[a * jb.L + k] : 0.0;
            const double cb = k < jb.len[b] ? jb.rows[b * jb.L + k] : 0.0;
            v = 0.5 * (cf + cb) * kALU * kALU * kALU;
        } else {
            if (dropped(jb.status[t], jb.status[t])) continue;
            const double cf = k < jb.len[t] ? jb.rows[t * jb.L + k] : 0.0;
            v = cf * kALU * kALU * kALU;
        }
        neumaier_add(s, cc, wr[t * ntemp + st] * v);
    }
    const long long o = ((long long)st * gridDim.y + c) * jb.L + k;
    psum[o] = s;
    pcomp[o] = cc;
}

// cf_out[st * L + k] = sum over chunks / nstar[st] (0 where nstar is 0: no trajectory carries
// weight at that temperature)
__global__ void __launch_bounds__(256) cf_reduce_weighted_final_kernel(int L, int nchunks, const double* __restrict__ nstar,
                                                                       const double* __restrict__ psum,
                                                                       const double* __restrict__ pcomp,
                                                                       double* __restrict__ cf_out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int st = blockIdx.y;
    if (k >= L) return;
    double s = 0.0, cc = 0.0;
    for (int c = 0; c < nchunks; ++c) {
        const long long o = ((long long)st * nchunks + c) * L + k;
        neumaier_add(s, cc, psum[o]);
        cc += pcomp[o];
    }
    cf_out[(long long)st * L + k] = nstar[st] > 0.0 ? (s + cc) / nstar[st] : 0.0;
}

// ------------------------------------------------------------------ host-side launchers
template <int SYS>
hipError_t launch_plain(const SysConst& sc, const JobDesc* jobs_dev, int njobs, long long total, double* ws,
                        hipStream_t st)
{
    const long long blocks = (total + 255) / 256;
    hipLaunchKernelGGL((plain_cf_kernel<SYS>), dim3((unsigned)blocks), dim3(256), 0, st, sc, jobs_dev, njobs, total,
                       ws);
    return hipGetLastError();
}

template <int SYS, typename P, int BS, int W>
hipError_t launch_pool(const SysConst& sc, const JobDesc* ring, const PoolCtl& ctl, const LaneSlots& sl, double* ws,
                       hipStream_t st)
{
    const unsigned grid = (unsigned)(sl.S / BS);
    hipLaunchKernelGGL((pool_cf_kernel<SYS, P, BS, W>), dim3(grid), dim3(BS), 0, st, sc, ring, ctl, sl, ws);
    return hipGetLastError();
}

}  // namespace hh

// explicit C++ entry points used by hh_host.cpp
namespace hh {
static int qp_of(int sys)
{
    return sys == SYS_HEAR ? SysTraits<SYS_HEAR>::QP : sys == SYS_CH4CO2 ? SysTraits<SYS_CH4CO2>::QP : SysTraits<SYS_CO2AR>::QP;
}

int arena_doubles_per_lane(int sys) { return kSlots * qp_of(sys); }

hipError_t run_plain(int sys, const SysConst& sc, const JobDesc* jobs_dev, int njobs, long long total, double* ws,
                     hipStream_t st)
{
    return sys == SYS_HEAR ? launch_plain<SYS_HEAR>(sc, jobs_dev, njobs, total, ws, st)
                           : launch_plain<SYS_CO2AR>(sc, jobs_dev, njobs, total, ws, st);
}

// the pool holds sl.S = grid * block lane slots; block is 64, 128 or 256 at two wavefronts per
// SIMD, 256 at one
hipError_t run_pool(int sys, bool mixed, int block, int waves, const SysConst& sc, const JobDesc* ring,
                    const PoolCtl& ctl, const LaneSlots& sl, double* ws, hipStream_t st)
{
#define HH_LAUNCH(S, PT) return launch_pool<S, PT, 256, 1>(sc, ring, ctl, sl, ws, st)
    // the one launch shape: 256-lane workgroups, one wavefront per SIMD
    if (block != 256 || waves != 1) return hipErrorInvalidValue;
    if (sys == SYS_CH4CO2) {
        if (mixed) HH_LAUNCH(SYS_CH4CO2, float);
        HH_LAUNCH(SYS_CH4CO2, double);
    }
    // He-Ar surfaces stay FP64 (see hear_dV); only the dipole follows P
    if (sys == SYS_HEAR) {
        if (mixed) HH_LAUNCH(SYS_HEAR, float);
        HH_LAUNCH(SYS_HEAR, double);
    }
    if (mixed) HH_LAUNCH(SYS_CO2AR, float);
    HH_LAUNCH(SYS_CO2AR, double);
#undef HH_LAUNCH
}

// lanes that found the queue empty advanced the head past the end; none of those indices was
// handed out, so the head goes back to the first unadmitted item
__global__ void clamp_next_kernel(unsigned long long* next, unsigned long long total)
{
    if (threadIdx.x == 0 && *next > total) *next = total;
}

hipError_t run_clamp_next(unsigned long long* next, unsigned long long total, hipStream_t st)
{
    hipLaunchKernelGGL(clamp_next_kernel, dim3(1), dim3(64), 0, st, next, total);
    return hipGetLastError();
}

hipError_t run_slots_clear(const LaneSlots& sl, hipStream_t st)
{
    hipLaunchKernelGGL(slots_clear_kernel, dim3((unsigned)((sl.S + 255) / 256)), dim3(256), 0, st, sl.i, sl.S);
    return hipGetLastError();
}

// src -> dst (dst is cleared first); `scratch` holds S/256 + 1 ints; the number of active
// slots lands in scratch[S/256]
hipError_t run_repack(int sys, const LaneSlots& src, const LaneSlots& dst, int* scratch, hipStream_t st)
{
    const unsigned nb = (unsigned)((src.S + 255) / 256);
    hipLaunchKernelGGL(repack_count_kernel, dim3(nb), dim3(256), 0, st, src.i, src.S, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(repack_scan_kernel, dim3(1), dim3(1024), 0, st, scratch, (int)nb, scratch + nb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = run_slots_clear(dst, st)) != hipSuccess) return e;
    const long long W = src.S / 64;
    if (sys == SYS_HEAR)
        hipLaunchKernelGGL(repack_move_kernel<SysTraits<SYS_HEAR>::QP>, dim3(nb), dim3(256), 0, st, src, dst, scratch, W);
    else if (sys == SYS_CH4CO2)
        hipLaunchKernelGGL(repack_move_kernel<SysTraits<SYS_CH4CO2>::QP>, dim3(nb), dim3(256), 0, st, src, dst, scratch, W);
    else
        hipLaunchKernelGGL(repack_move_kernel<SysTraits<SYS_CO2AR>::QP>, dim3(nb), dim3(256), 0, st, src, dst, scratch, W);
    return hipGetLastError();
}

hipError_t run_eval(int sys, const SysConst& sc, int what, const double* y, long long n, double* out, hipStream_t st)
{
    const unsigned g = (unsigned)((n + 255) / 256);
    if (sys == SYS_HEAR)
        hipLaunchKernelGGL(eval_kernel<SYS_HEAR>, dim3(g), dim3(256), 0, st, sc, what, y, n, out);
    else if (sys == SYS_CH4CO2)
        hipLaunchKernelGGL(eval_kernel<SYS_CH4CO2>, dim3(g), dim3(256), 0, st, sc, what, y, n, out);
    else
        hipLaunchKernelGGL(eval_kernel<SYS_CO2AR>, dim3(g), dim3(256), 0, st, sc, what, y, n, out);
    return hipGetLastError();
}

// keys (of kind `what`) and item numbers of job jb's items, written at
// keys[0..n_items) and idx[0..n_items) with item numbers counted from idx_offset
hipError_t run_order_keys(const JobDesc& jb, double mu, int what, float* keys, int* idx, int idx_offset,
                          hipStream_t st)
{
    const long long n = jb.n_items;
    hipLaunchKernelGGL(order_keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, jb.ic, jb.ntraj,
                       jb.halves, mu, jb.Rcut, jb.dt, jb.L, what, keys, idx, idx_offset);
    return hipGetLastError();
}

hipError_t run_order_front(const int* order, long long n, long long n_front, float* keys, hipStream_t st)
{
    hipLaunchKernelGGL(order_front_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, order, n, n_front,
                       keys);
    return hipGetLastError();
}

size_t slots_bytes(int sys, long long S)
{
    const int N = qp_of(sys);
    return (size_t)S * ((2 * N + 8) * sizeof(double) + 6 * sizeof(int) + sizeof(long long));
}

LaneSlots slots_view(int sys, void* base, long long S)
{
    const int N = qp_of(sys);
    LaneSlots sl;
    sl.S = S;
    sl.d = static_cast<double*>(base);
    sl.local = reinterpret_cast<long long*>(sl.d + (size_t)(2 * N + 8) * S);
    sl.i = reinterpret_cast<int*>(sl.local + S);
    return sl;
}

hipError_t run_reduce_weighted(const JobDesc& jb, int ntemp, const double* wr, const double* nstar, double* psum,
                               double* pcomp, long long chunk, int nchunks, double* cf_out, hipStream_t st)
{
    dim3 g1((jb.L + 255) / 256, nchunks, ntemp);
    hipLaunchKernelGGL(cf_reduce_weighted_partial_kernel, g1, dim3(256), 0, st, jb, chunk, ntemp, wr, psum, pcomp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cf_reduce_weighted_final_kernel, dim3((jb.L + 255) / 256, ntemp), dim3(256), 0, st, jb.L,
                       nchunks, nstar, psum, pcomp, cf_out);
    return hipGetLastError();
}

hipError_t run_reduce(const JobDesc& jb, double ppf, double* psum, double* pcomp, long long chunk, int nchunks,
                      double* cf_out, hipStream_t st)
{
    dim3 g1((jb.L + 255) / 256, nchunks);
    hipLaunchKernelGGL(cf_reduce_partial_kernel, g1, dim3(256), 0, st, jb, chunk, ppf, psum, pcomp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cf_reduce_final_kernel, dim3((jb.L + 255) / 256), dim3(256), 0, st, jb.L, nchunks, jb.ntraj,
                       psum, pcomp, cf_out);
    return hipGetLastError();
}
}  // namespace hh
