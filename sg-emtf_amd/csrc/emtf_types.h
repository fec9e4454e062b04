// emtf_types.h -- host/device descriptors of the island engine.
#pragma once
#include <stdint.h>

namespace emtf {

// 0..7: this engine's binary32 forms; 8..13: the reference's own formulas and constants
// (evaluation(), SG-EMTF_DE.cu:834-1069).  12 runs the Schwefel body with the reference's
// constant in the table; 13 the reference Griewank body with +inf divisors (the GA copy's
// disabled product, SG-EMTF.cu:1029: every factor cos(z / inf) = 1).
enum Func : int32_t {
    F_SPHERE = 0, F_ROSENBROCK = 1, F_ACKLEY = 2, F_RASTRIGIN = 3,
    F_GRIEWANK = 4, F_WEIERSTRASS = 5, F_SCHWEFEL = 6, F_ELLIPTIC = 7,
    F_REF_ACKLEY = 8, F_REF_RASTRIGIN = 9, F_REF_GRIEWANK = 10, F_REF_WEIERSTRASS = 11,
    F_REF_SCHWEFEL = 12, F_REF_GRIEWANK_GA = 13
};
enum Algo : int32_t { A_DE_RAND_1_BIN = 0, A_DE_CURRENT_1_BIN = 1, A_GA_SBX_PM = 2 };

// One task as the kernels see it.  All offsets are in elements.
struct TaskDev {
    int32_t func, dim, pop, xform;
    float lo, hi;
    // life on the handle's generation clock: the task produces generations gen_base + 1 ..
    // gen_end (0 and INT32_MAX unless admitted online; a task not yet admitted has
    // gen_base = INT32_MAX); its own generation counter, which addresses its draws, is
    // the clock minus gen_base
    int32_t gen_base, gen_end;
    int64_t x_off;    // population X[pop][dim], row-major
    int64_t fit_off;  // fitness[pop]
    int64_t q_off;    // transform matrix Q[dim][dim], row-major (identity / diagonal materialised)
    int64_t v_off;    // shift[dim] and per-dimension table[dim]
    int64_t mb_off;   // elite mailbox [transfer_count][dim] inside one mailbox buffer
    int64_t scr_off;  // scratch [pop][row stride] (wide tasks: population rows, then trial rows)
    int64_t qw_off;   // wide tasks: padded Q image [32 NT][row stride]
};

struct EngineParams {
    int32_t algo;
    float F, CR, eta_c, eta_m, pm;
    int32_t n_islands, exchange_interval, migrate_interval, migrate_count;
    int32_t transfer_interval, transfer_count;
    float transfer_cr;
    uint32_t seed_lo, seed_hi;
    uint32_t cr_threshold;  // floor(CR 2^16) clamped to [0, 65536]: keep the parent iff u16 >= it
    int32_t polish_interval, polish_steps, polish_topk;  // gradient polish of elites (0 = off)
};

struct EpochArgs {
    const TaskDev* tasks;
    int32_t n_tasks;
    int32_t g0, g1;     // generations g0+1 .. g1 are produced
    int32_t init;       // 1: create and evaluate generation 0 (g0 == g1 == 0)
    float* X;
    float* fit;
    const float* Q;
    const float* shift;
    const float* tab;
    float* mailbox;
    int64_t mailbox_stride;  // elements per mailbox buffer
    float* scratch;
    const float* Qw;   // padded Q images of the wide tasks
    EngineParams p;
    int32_t publish;       // write the task's elites to the mailbox after generation g1
    // Replayed epochs (hipGraph): when set, the epoch starts at *gen_dev instead of g0 and
    // keeps its length g1 - g0 (kernel arguments are frozen at capture; the counter is not)
    const int32_t* gen_dev;
    int32_t polish_waves;  // polish kernel: waves working on elites (1..4)
    int32_t polish_cpe;    // polish kernel: MFMA columns per elite (power of two, 1..32)
    // Transfer donors (online admission): donors[0 .. *n_donors) are the ids, ascending, of
    // the tasks that ran up to the transfer boundary g0 and run on after it; null: every
    // task of the handle
    const int32_t* donors;
    const int32_t* n_donors;
};

// The persistent schedule: one launch runs generations G0+1 .. G1 of every task as a
// sequence of epochs cut at the transfer boundaries.  Items (epoch e, task order[i]),
// e-major, are taken from a device counter; an item waits (thread 0, device-scope acquire
// loads, a bounded wait) for its task's previous epoch and, at a boundary, for its donor's
// previous epoch (whose end published the elites it reads); an item that publishes first
// waits until every item of the previous epoch has read the mailbox buffer it overwrites.
struct PersistArgs {
    int32_t G0, G1;        // generations G0+1 .. G1
    int32_t n_epochs;      // epochs in this launch
    int32_t* next;         // item counter (starts at 0)
    int32_t* err;          // set to 1 when a wait ran out (starts at 0)
    int32_t* stamp;        // [n_tasks] epochs of this launch each task has finished (start 0)
    int32_t* reads;        // [n_epochs] items of each epoch past their mailbox read (start 0)
    int64_t wait_ticks;    // bound of one wait in wall-clock ticks (100 MHz)
};

// Stand-alone objective / adjoint evaluation of n points (device pointers).
struct ObjArgs {
    int32_t func, dim, n, has_rot;
    const float* x;      // [n][dim]
    const float* rot;    // [dim][dim] or null (identity)
    const float* shift;  // [dim] or null (zero)
    const float* gf;     // [n] upstream gradient (adjoint)
    float* f;            // [n] objective (forward)
    float* gx;           // [n][dim] gradient (adjoint)
    float tab[256];      // per-dimension table (Elliptic weights / Griewank scales)
};

}  // namespace emtf
