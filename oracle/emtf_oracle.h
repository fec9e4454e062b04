/*
 * emtf_oracle.h -- CPU oracle for the island-model evolutionary-multitasking
 * engine.  TEST INFRASTRUCTURE ONLY: the product never includes, links or
 * calls anything under oracle/.  Only tests/, the smoke check and the CPU
 * baseline leg of the benchmark load liboracle.so, and only as the checker or
 * as the timed CPU baseline.
 *
 * The oracle is a plain-C restatement of the reference operators
 * (proposed model/SG-EMTF.cu, SG-EMTF_DE.cu), with the reference's cuRAND
 * XORWOW streams replaced by counter-addressed Philox4x32-10 draws and its
 * CUDA libm calls replaced by the FP32 function definitions below, so that a
 * GPU implementation following the same arithmetic reproduces every bit.
 *
 * ---------------------------------------------------------------------------
 * Arithmetic rules (shared contract with the device code)
 *   - IEEE binary32, round-to-nearest-even, no contraction: every a*b+c is two
 *     roundings unless written fmaf(a,b,c).  Build with -ffp-contract=off.
 *   - Division and sqrtf are correctly rounded.
 *   - The rotation z = Q w is the fmaf chain acc = fmaf(Q[j][d], w[d], acc)
 *     for d = 0 .. Dp-1 (Dp = D rounded up to even, padding zeros), starting
 *     from +0.  This is exactly what a sequence of 32x32x2 FP32 MFMA steps
 *     computes on gfx950.
 *   - Per-dimension terms t_j are summed in two accumulators: terms whose
 *     block index j/4 is even go to A, odd to B, each in increasing j, both
 *     starting at +0; the sum is A + B.
 *
 * Random draws: Philox4x32-10 (Salmon et al., SC'11) with
 *   key     = { seed & 0xffffffff, seed >> 32 }
 *   counter = { block, purpose<<24 | sub<<16 | index, generation, task }
 * Every draw is addressed by (task, generation, purpose, index, sub, word),
 * never by a thread id, so any thread mapping consumes identical numbers.
 * ---------------------------------------------------------------------------
 */
#ifndef EMTF_ORACLE_H
#define EMTF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 0..7: the objectives in this engine's binary32 forms (cancellation-free, textbook
 * constants).  8..13: the reference's own formulas and constants (the evaluation() of
 * SG-EMTF_DE.cu:834-1069; 13 is the GA copy's Griewank, SG-EMTF.cu:1029, whose product is
 * disabled); 12 and 13 share the bodies of 6 and 10, their tables differ. */
enum {
    OR_FUNC_SPHERE = 0, OR_FUNC_ROSENBROCK = 1, OR_FUNC_ACKLEY = 2, OR_FUNC_RASTRIGIN = 3,
    OR_FUNC_GRIEWANK = 4, OR_FUNC_WEIERSTRASS = 5, OR_FUNC_SCHWEFEL = 6, OR_FUNC_ELLIPTIC = 7,
    OR_FUNC_REF_ACKLEY = 8, OR_FUNC_REF_RASTRIGIN = 9, OR_FUNC_REF_GRIEWANK = 10, OR_FUNC_REF_WEIERSTRASS = 11,
    OR_FUNC_REF_SCHWEFEL = 12, OR_FUNC_REF_GRIEWANK_GA = 13,
    OR_NUM_FUNCS = 14
};
enum { OR_XFORM_SHIFT = 0, OR_XFORM_ROTATE = 1, OR_XFORM_DIAG = 2 };
enum { OR_ALGO_DE_RAND_1_BIN = 0, OR_ALGO_DE_CURRENT_1_BIN = 1, OR_ALGO_GA_SBX_PM = 2 };
enum {
    OR_P_INIT = 1, OR_P_DE_IDX = 2, OR_P_DE_CR = 3, OR_P_GA_MATE = 4, OR_P_GA_SBX = 5,
    OR_P_GA_PM = 6, OR_P_MIG = 7, OR_P_XFER = 8, OR_P_ISLX = 9, OR_P_CV_INIT = 16, OR_P_CV_GEN = 17
};

/* One continuous task.  z = Q (x - o) (ROTATE), z = x - o (SHIFT),
 * z = M * (x - o) (DIAG, reference transform z = M (x + b) with o = -b). */
typedef struct {
    int32_t func, dim, pop, xform;
    float lo, hi;
    const float* rot;    /* dim*dim row-major Q[j][d]   (ROTATE) */
    const float* shift;  /* dim                          (all)    */
    const float* diag;   /* dim                          (DIAG)   */
} or_task;

typedef struct {
    int32_t algo;
    float F, CR;                  /* DE */
    float eta_c, eta_m, pm;       /* GA: SBX index, polynomial-mutation index and rate */
    int32_t n_islands;            /* sub-populations per task (mates drawn inside one) */
    int32_t exchange_interval;    /* cross-island resampling period (0 = off) */
    int32_t migrate_interval;     /* intra-task island migration period (0 = off) */
    int32_t migrate_count;
    int32_t transfer_interval;    /* inter-task transfer period (0 = off) */
    int32_t transfer_count;
    float transfer_cr;            /* keep own gene if u <= transfer_cr (0 = take all) */
    uint64_t seed;
    int32_t polish_interval;      /* gradient polish period (0 = off) */
    int32_t polish_steps;         /* Armijo steps per elite */
    int32_t polish_topk;          /* elites polished per task */
} or_params;

/* ---- Philox ------------------------------------------------------------ */
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_draw_block(uint64_t seed, uint32_t task, uint32_t gen, uint32_t purpose,
                   uint32_t index, uint32_t sub, uint32_t block, uint32_t out[4]);

/* ---- FP32 function definitions (bit-exact contract) ---------------------- */
float or_sinpi_sq(float v);      /* sin^2(pi v) */
float or_sin2pi(float v);         /* sin(2 pi v) */
float or_expm1(float x);          /* e^x - 1, x <= 88 */
float or_exp2(float y);
float or_log2(float x);
float or_pow(float x, float y);   /* x >= 0 */
float or_weier_term(float z);     /* sum_k 2^(1-k) sin^2(pi 3^k z), k = 0..20 */
float or_cospi(float y);          /* cos(pi y), any finite y (exact reduction) */
float or_exp(float x);            /* e^x = expm1(x) + 1 */
float or_weier_ref_term(float z); /* sum_k 0.5^k cospi(2 (z + 0.5) 3^k) in binary32, k = 0..20 */
float or_weier_ref_w(void);       /* the reference's Weierstrass constant (PI = 3.1415926) */

/* Per-dimension coefficient tables: Elliptic weights, Griewank scales, the Schwefel
 * constant, the reference Griewank divisors sqrt(j+1) (+inf for the GA copy: each
 * factor is then cos(0) = 1) and the reference Weierstrass constant. */
void or_func_table(int32_t func, int32_t dim, float* table);

/* Objective from z (contract arithmetic), table from or_func_table. */
float or_finish32(const or_task* t, const float* z, const float* tab);
/* FP32 evaluation of n points x[n][dim] (contract arithmetic). */
void or_eval32(const or_task* t, int32_t n, const float* x, float* f);
/* FP64 textbook evaluation (independent formulas, libm). */
void or_eval64(const or_task* t, int32_t n, const double* x, double* f);
/* FP32 gradient (contract arithmetic of the device adjoint). */
void or_grad32(const or_task* t, int32_t n, const float* x, float* gx);
/* FP64 gradient of the textbook objective at x[n][dim]. */
void or_grad64(const or_task* t, int32_t n, const double* x, double* g);

/* ---- Engine -------------------------------------------------------------- */
typedef struct or_engine or_engine;
or_engine* or_create(const or_task* tasks, int32_t n_tasks, const or_params* p);
/* Online admission: or_create_deferred starts no task; or_admit starts tasks first ..
 * first + count - 1 at the engine's clock (a transfer boundary) for `generations`
 * generations, each with its own generation counter from 0.  At a transfer boundary the
 * donors are the tasks that ran up to it and run on after it. */
or_engine* or_create_deferred(const or_task* tasks, int32_t n_tasks, const or_params* p);
int32_t or_admit(or_engine* e, int32_t first, int32_t count, int32_t generations);
/* The donor `task` draws at the transfer boundary `clock` under the current admissions
 * (-1: it receives nothing there). */
int32_t or_transfer_donor(const or_engine* e, int32_t task, int32_t clock);
void or_destroy(or_engine* e);
/* Advance every task by n generations.  n_threads <= 0: OpenMP default. */
void or_run(or_engine* e, int32_t n_gens, int32_t n_threads);
int32_t or_generation(const or_engine* e);
void or_get_population(const or_engine* e, int32_t task, float* x, float* fit);
void or_set_population(or_engine* e, int32_t task, const float* x, const float* fit);
void or_get_best(const or_engine* e, float* fit, float* x_best, int32_t x_stride);
/* Single phases on the current state (tests pin them one at a time): the cross-island
 * exchange of generation gen, the island migration, and one inter-task transfer event
 * (every task's elites snapshotted, then every task receives its partner's). */
void or_phase_exchange(or_engine* e, int32_t task, int32_t gen);
void or_phase_migrate(or_engine* e, int32_t task);
void or_phase_transfer(or_engine* e, int32_t gen);

/* ---- CVRP (exact integers) -------------------------------------------------- */
typedef struct {
    int32_t n;               /* customers 1..n, depot 0; n <= 255 */
    int32_t capacity;
    int32_t pop;
    const int32_t* xy;       /* (n+1) x 2 integer coordinates, row 0 = depot */
    const int32_t* demand;   /* n+1, demand[0] = 0, each >= 1 */
} or_cvrp_instance;

typedef struct {
    float p_swap, p_2opt;    /* mutation: swap with p_swap, else 2-opt with p_2opt */
    int32_t n_islands, migrate_interval, migrate_count;
    uint64_t seed;
} or_cvrp_params;

typedef struct or_cvrp or_cvrp;
int32_t or_cvrp_dist(int32_t x0, int32_t y0, int32_t x1, int32_t y1);
int32_t or_cvrp_split(const uint8_t* tour, int n, const uint8_t* dist, const uint8_t* dem, int cap);
void or_cvrp_child(const uint8_t* p1, const uint8_t* p2, int n, int a, int b, int mut, int c, int d,
                   uint8_t* child);
or_cvrp* or_cvrp_create(const or_cvrp_instance* inst, int32_t n_inst, const or_cvrp_params* p);
void or_cvrp_run(or_cvrp* e, int32_t gens, int32_t n_threads);
void or_cvrp_get(const or_cvrp* e, int32_t inst, uint8_t* tours, int32_t* cost);
void or_cvrp_set(or_cvrp* e, int32_t inst, const uint8_t* tours, const int32_t* cost);
void or_cvrp_phase_migrate(or_cvrp* e, int32_t inst);
void or_cvrp_destroy(or_cvrp* e);

#ifdef __cplusplus
}
#endif
#endif
