// Structures shared by the host library and the device kernels of the trajectory engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hh {

// Per-work-item outcome codes (an item is one half of one trajectory).
enum ItemStatus : int {
    ITEM_OK = 0,
    ITEM_FAIL_DIPOLE_NAN = 1,   // src/hawaii.c:2181-2190
    ITEM_FAIL_CF_HUGE = 2,      // src/hawaii.c:2212-2219
    ITEM_FAIL_CF_JUMP = 3,      // src/hawaii.c:2221-2236
    ITEM_FAIL_STEPS = 4,        // too many steps between two samples (CVODE mxsteps analogue)
    ITEM_FAIL_STEPSIZE = 5,     // step size underflow
    ITEM_ORBIT = 6,             // mixed precision: handed over for integration in FP64 (an orbiting pair);
                                // replaced by the FP64 run's own status before the job is reduced
};

// Values per sample of a job's rows for its record mode: 0 mu(0).mu(t_k); 1, 2 the dipole vector
// (Zimmermann's trick, PR/mu); 3 the eight lab coordinates of a CH4-CO2 sample (its dipole is
// evaluated afterwards by the batched dipole kernel, which overwrites the first three with mu)
__host__ __device__ __forceinline__ int row_width(int record) { return record == 3 ? 8 : (record ? 3 : 1); }

__host__ __device__ __forceinline__ bool dipole_failure(int st)
{
    return st == ITEM_FAIL_DIPOLE_NAN || st == ITEM_FAIL_CF_HUGE || st == ITEM_FAIL_CF_JUMP;
}

// A trajectory contributes zero when either half failed a dipole check (crln stays zeroed,
// src/hawaii.c:2183-2237, and src/hawaii.c:1957-2065 with the Zimmermann trick) or when its
// last half ended on the step limit: correlation_eval then returns CVode's CV_TOO_MUCH_WORK
// (-1) and the caller skips it (src/hawaii.c:3001).  It still counts in ntraj.
__host__ __device__ __forceinline__ bool dropped(int first, int last)
{
    return dipole_failure(first) || dipole_failure(last) || last == ITEM_FAIL_STEPS;
}

// Scheduling keys of a job's items (order_keys_kernel): the straight-line transit time through the
// Rcut sphere, the squared impact parameter, and the transit time with a marked front tier.
enum OrderKey : int { ORDER_TRANSIT_KEY = 0, ORDER_IMPACT_KEY = 1, ORDER_TIERED_KEY = 2 };

// hh_cf_to_sf_noalloc: the largest transform length its static work array holds (two
// 2n-point complex FP64 sequences: 4 MiB per device)
constexpr long kCfToSfNoAllocMaxN = 65536;

// One correlation-function job: a batch of trajectories sharing system, temperature-free
// integration parameters and output buffers.  All pointers are device pointers.
struct JobDesc {
    const double* ic;        // initial conditions, SoA [QP][ntraj]
    double* rows;            // per-item samples, [n_items][L] or [n_items][3][L] (record), only [0, len) written
    int* len;                // per-item number of written samples
    int* status;             // per-item ItemStatus
    int* nrhs;               // per-item right-hand-side evaluations
    float* e_drift;          // per-item |H(last sample) - H(0)|
    long long ntraj;
    long long item_begin;    // first global item index of this job
    long long n_items;       // ntraj * halves
    int* done;               // finished-item counter of the job (nullptr: not counted)
    const int* order;        // admission group's order: the item run at queue position g is
    long long order_base;    //   order_base + order[g - order_base] (nullptr: the item g itself)
    int halves;              // 1: forward only, 2: forward + momentum-inverted backward
    int L;                   // MaxTrajectoryLength
    int max_steps;           // max accepted steps between two samples
    int record;              // 0: rows hold mu(0).mu(t_k) [L]; 1: the dipole series mu_x, mu_y, mu_z [3][L]
                             // (Zimmermann); 2: the same series for PR/mu, dipole check only, no CF reduce
    double dt, Rcut, rtol, atol;
    // Item map of an orbit job: its item k is item map[k] of the parent job whose buffers (ic, rows,
    // len, status, nrhs, e_drift) it shares; nullptr: item k is item k.
    const int* map;
    // Mixed precision: an item whose pair reaches an outer radial turning point (pR from > 0 to
    // <= 0, the pair bound in an orbiting complex) ends with ITEM_ORBIT and its index is appended
    // to orbit_items[atomicAdd(orbit_count, 1)]; nullptr: no hand-over.
    int* orbit_items;
    int* orbit_count;
    // an orbit job: the parent job's counter of the right-hand sides spent in FP64
    unsigned long long* orbit_nrhs;
    // record mode 3: the CF rows [n_items][L] the dipole pass writes (what the reduction reads),
    // and whether that pass evaluates the dipole in FP32
    double* cf_rows;
    int f32_dipole;
};

struct QueueState {
    unsigned long long next;     // next global item to hand out
    unsigned long long done;     // items finished
};

// Resident lane pool of the batched engine: the integrator state of every lane slot,
// structure-of-arrays with stride S so a wavefront's save / restore is coalesced.  A slot
// whose job field is negative holds no item.
//   d: y[N], k0[N] (= f(y), first-same-as-last), t, h, E0, mu0[3], prev, curr -> (2N + 8) x S
//   i: job (ring index), k, steps, total_steps, nrhs, reject                  -> 6 x S
//   local: item index inside its job                                           -> S
struct LaneSlots {
    double* d;
    int* i;
    long long* local;
    long long S;
};

// One slice of the pool kernel: items [0, total) of the class's queue have been admitted;
// jobs with sequence numbers [first_seq, last_seq] live in ring slots seq % ring.  Lanes stop
// at the first step boundary after `ticks` wall-clock ticks (s_memrealtime) of the slice.
struct PoolCtl {
    unsigned long long* next;
    unsigned long long total;
    long long first_seq, last_seq;
    long long ticks;
    int ring;
    int pad;
};

// Rejection-sampler M0 (calculate_M0, src/hawaii.c:1423-1491)
struct M0Params {
    double T, Rmin, Rmax, pesmin;
    int pair_state;
    int pad;
};

// Per-stream accumulators of a (possibly sliced) M0 run: running mean m and sum of squared
// deviations q (Welford, src/hawaii.c:1470-1480), accepted points n, candidates drawn, MT
// read position; `finished` counts the streams that reached their quota or candidate limit.
struct M0Acc {
    double* m;
    double* q;
    long* n;
    long* cand;
    int* pos;
    int* finished;
};

}  // namespace hh
