// Structures shared by the host library and the device kernels of the trajectory engine.
#pragma once
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
};

// One correlation-function job: a batch of trajectories sharing system, temperature-free
// integration parameters and output buffers.  All pointers are device pointers.
struct JobDesc {
    const double* ic;        // initial conditions, SoA [QP][ntraj]
    double* rows;            // per-item CF samples mu(0).mu(t_k), [n_items][L] (only [0, len) written)
    int* len;                // per-item number of written samples
    int* status;             // per-item ItemStatus
    int* nrhs;               // per-item right-hand-side evaluations
    float* e_drift;          // per-item |H(last sample) - H(0)|
    long long ntraj;
    long long item_begin;    // first global item index of this job
    long long n_items;       // ntraj * halves
    int halves;              // 1: forward only, 2: forward + momentum-inverted backward
    int L;                   // MaxTrajectoryLength
    int max_steps;           // max steps between two samples
    int pad;
    double dt, Rcut, rtol, atol;
};

struct QueueState {
    unsigned long long next;     // next global item to hand out
    unsigned long long done;     // items finished
};

}  // namespace hh
