/*
 * emtf.h -- C ABI of the MI355X island-model evolutionary-multitasking engine.
 *
 * The reference (MZT-DW/SG-EMTF, proposed model/) has no library interface:
 * its "API" is a set of compile-time knobs and kernel launches made from one
 * host thread per task.  Each entry point below names the reference code it
 * replaces.  All functions return 0 (EMTF_OK) or a negative EMTF_E_* code;
 * emtf_last_error() gives a message for the calling thread.  Host pointers are
 * copied; device buffers are owned by the handle.  A handle is used by one host
 * thread at a time; independent handles may run concurrently on different
 * streams.  Plain C types only.
 */
#ifndef EMTF_H
#define EMTF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMTF_OK 0
#define EMTF_E_INVALID (-1)     /* a descriptor or argument is out of range */
#define EMTF_E_DEVICE (-2)      /* the HIP runtime reported an error */
#define EMTF_E_NOMEM (-3)       /* device allocation failed */
#define EMTF_E_UNSUPPORTED (-4) /* shape outside what the built kernels cover */

/* Objectives: ids 0..6 are the reference's evaluation() cases
 * (proposed model/SG-EMTF_DE.cu:857-1061); 7 is the CEC Elliptic. */
enum {
    EMTF_SPHERE = 0, EMTF_ROSENBROCK = 1, EMTF_ACKLEY = 2, EMTF_RASTRIGIN = 3,
    EMTF_GRIEWANK = 4, EMTF_WEIERSTRASS = 5, EMTF_SCHWEFEL = 6, EMTF_ELLIPTIC = 7
};
/* z = x - o | z = Q (x - o) | z = M (x - o)  (reference: z = M (x + b), SG-EMTF_DE.cu:859) */
enum { EMTF_SHIFT = 0, EMTF_ROTATE = 1, EMTF_DIAG = 2 };
/* DE/rand/1/bin | DE/current/1/bin (reference CrossOver+Mutation, SG-EMTF_DE.cu:474-677)
 * | SBX + polynomial mutation (reference GA, SG-EMTF.cu:548-831) */
enum { EMTF_DE_RAND_1_BIN = 0, EMTF_DE_CURRENT_1_BIN = 1, EMTF_GA_SBX_PM = 2 };

/* One continuous task.  Replaces the per-task globals task_type / M / b /
 * Range (SG-EMTF.cu:92-124, rangeInit :197-214, para_init :181-195). */
typedef struct emtf_task_desc {
    int32_t func;        /* EMTF_SPHERE .. EMTF_ELLIPTIC */
    int32_t dim;         /* 1 .. 256 */
    int32_t pop;         /* individuals, a multiple of n_islands */
    int32_t xform;       /* EMTF_SHIFT / EMTF_ROTATE / EMTF_DIAG */
    float lo, hi;        /* search box, every dimension */
    const float* rot;    /* dim*dim row-major Q (EMTF_ROTATE), else NULL */
    const float* shift;  /* dim, o (NULL = 0) */
    const float* diag;   /* dim, M (EMTF_DIAG), else NULL */
} emtf_task_desc;

/* Engine knobs.  Replaces the #define block SG-EMTF.cu:28-68 / SG-EMTF_DE.cu:23-59:
 * F, CROSSOVER_RATE, ETA, DM, ISLAND_NUM/INDIV_PERBLOCK, INTERVAL_MIGRA,
 * MIGRA_NUM, INTERVAL_TRANSFER, TRANSFER_NUM. */
typedef struct emtf_params {
    int32_t algo;               /* EMTF_DE_RAND_1_BIN / EMTF_DE_CURRENT_1_BIN / EMTF_GA_SBX_PM */
    float F, CR;                /* DE scale and binomial crossover rate */
    float eta_c, eta_m, pm;     /* GA: SBX index, mutation index, per-gene mutation rate */
    int32_t n_islands;          /* sub-populations per task */
    int32_t exchange_interval;  /* cross-island resampling period in generations (0 = off) */
    int32_t migrate_interval;   /* intra-task island migration period (0 = off) */
    int32_t migrate_count;      /* individuals per island pair, <= island size / 2 */
    int32_t transfer_interval;  /* inter-task elite transfer period (0 = off) */
    int32_t transfer_count;     /* elites per transfer, <= pop / 2 */
    float transfer_cr;          /* keep own gene when u <= transfer_cr; 0 takes every donor gene */
    uint64_t seed;
    int32_t polish_interval;    /* gradient polish period in generations (0 = off); no reference counterpart */
    int32_t polish_steps;       /* Armijo steps per polished elite */
    int32_t polish_topk;        /* elites polished per task */
    int32_t schedule;           /* EMTF_SCHEDULE_*; no reference counterpart */
    int32_t tune;               /* 0: autotune kernel variants and stream balance from measured
                                   epoch times (results are the same bits either way; the
                                   decisions are kept for later creates of a batch with the same
                                   launch units: in memory for the process, and in the file the
                                   environment variable EMTF_TUNE_CACHE names, if set); 1: off;
                                   2: the first alternative variant where one exists (the wide
                                   kernel, or its 512-thread build); 3: the wide kernel with Q in
                                   LDS where it fits (2 and 3 for testing) */
    int32_t formulas;           /* EMTF_FORMULAS_*: which formulas evaluate the objectives */
    int32_t admission;          /* EMTF_ADMIT_*: when tasks start */
} emtf_params;

/* Objective formulas.  TEXTBOOK: the functions in cancellation-free binary32 forms with
 * exact constants (zero at every optimum).  REFERENCE_DE: the reference's own formulas and
 * constants -- evaluation() of SG-EMTF_DE.cu:834-1069: Ackley with E = 2.718282, Rastrigin
 * as z^2 - 10 cospi(2z) + 10, Griewank's product of cosf(z / sqrtf(j+1)), Weierstrass with
 * the binary32-rounded cospif arguments and the constant computed from PI = 3.1415926
 * (SG-EMTF.cu:354-361), Schwefel with 418.9829.  REFERENCE_GA: the same, as the GA copy
 * (SG-EMTF.cu:898-1111) has it: Griewank without the product (SG-EMTF.cu:1029).  Sphere,
 * Rosenbrock and Elliptic are the same in every mode.  Gradient polish needs TEXTBOOK. */
enum { EMTF_FORMULAS_TEXTBOOK = 0, EMTF_FORMULAS_REFERENCE_DE = 1, EMTF_FORMULAS_REFERENCE_GA = 2 };

/* How emtf_run launches work.  Work is cut into epochs that end at every transfer
 * boundary, the kernel boundary ordering the elite mailbox.  EPOCHS: each epoch's kernels
 * (one per launch unit, forked over up to four streams and joined) are launched directly.
 * AUTO: the same epochs, replayed from a hipGraph that the handle captures once (and again
 * when the epoch length changes): epochs of one whole transfer interval (or of any length
 * without transfer) whose end carries no trace point or polish; the kernels read their
 * first generation from a device counter that the graph advances.  Same results either way.
 * PERSISTENT: one launch per call (per stretch between trace points, at most 1024 epochs
 * each) of one workgroup per (epoch, task) item; each workgroup takes the next item, in the
 * order the workgroups start, from a device counter, and waits, on the device, for its
 * task's and its donor's previous epoch and, before it publishes, for the previous epoch's
 * mailbox readers; every wait is bounded (10 s), and one that runs out makes emtf_run return
 * EMTF_E_DEVICE.  emtf_run then waits for the stream.
 * Only for handles whose tasks all fall in one LDS-kernel bucket (one row geometry, the
 * population in LDS), without polish or online admission (else EMTF_E_UNSUPPORTED at
 * create); the autotuner and the graph are off.  Same results as the other schedules. */
enum { EMTF_SCHEDULE_AUTO = 0, EMTF_SCHEDULE_EPOCHS = 1, EMTF_SCHEDULE_PERSISTENT = 2 };

/* When tasks start.  AT_CREATE: every task starts at emtf_create.  ONLINE: none does;
 * emtf_admit starts batches of them while earlier ones run (the reference's online driver,
 * SG-EMTF.cu:1842-1958).  The handle's generation counter is then a clock: an admitted task
 * runs `generations` generations from the clock value at its admission, with its own
 * generation counter (the one its draws are addressed by) starting at 0, and transfers draw
 * their donor among the tasks that ran up to the transfer boundary and run on after it --
 * across batches, as the reference draws among every live task (SG-EMTF.cu:1696-1699).
 * ONLINE runs the epochs directly (no graph replay), without autotuning and without polish. */
enum { EMTF_ADMIT_AT_CREATE = 0, EMTF_ADMIT_ONLINE = 1 };

typedef struct emtf_handle emtf_handle;

/* Allocate device state, create generation 0 and evaluate it.
 * Replaces initialization() (SG-EMTF.cu:293-380) + pop_init/evaluate (:146, :1428). */
int emtf_create(const emtf_task_desc* tasks, int32_t n_tasks, const emtf_params* params, emtf_handle** out);

/* Advance every task by `generations` (asynchronous on `stream`, a hipStream_t;
 * NULL = default stream; a run on another stream than the previous call's is ordered
 * after that call's work).  Replaces iter()'s loop of GeneticOpera launches and
 * the knowledgeTransfer handshake (SG-EMTF.cu:1646-1745, kernel :1289). */
int emtf_run(emtf_handle* h, int32_t generations, void* stream);

/* ONLINE admission: start tasks first .. first + count - 1 (not admitted before) at the
 * handle's clock, which must be a transfer boundary when transfer is on; they are created
 * and evaluated on `stream` and run the next `generations` generations of the clock.
 * Replaces the per-task host thread start of the reference's online driver
 * (SG-EMTF.cu:1646-1660, :1890-1897). */
int emtf_admit(emtf_handle* h, int32_t first, int32_t count, int32_t generations, void* stream);

/* Wait for the work queued by emtf_run. */
int emtf_sync(emtf_handle* h);

/* Generations completed so far. */
int emtf_generation(const emtf_handle* h, int32_t* generation);

/* Copy one task's population x[pop][dim] and fitness[pop] to host memory
 * (synchronises).  Replaces popTransfer (SG-EMTF.cu:216-229). */
int emtf_get_population(emtf_handle* h, int32_t task, float* x, float* fitness);

/* Overwrite one task's population and fitness from host memory (synchronises). */
int emtf_set_population(emtf_handle* h, int32_t task, const float* x, const float* fitness);

/* Best fitness of every task and its genome (x_best rows of stride x_stride >= max dim).
 * Replaces popSort + messageTransfer (SG-EMTF.cu:462-486, :1601-1615). */
int emtf_get_best(emtf_handle* h, float* fitness, float* x_best, int32_t x_stride);

/* The k best individuals of every task, ranked on the device by (fitness, index) with NaN
 * last: fitness[T][k] and x[T][k][x_stride] (x_stride >= max dim; rows zero past each task's
 * dim; either output may be NULL; k <= every task's pop; synchronises).  Replaces popSort +
 * messageTransfer with BESTNUM_PERTASK = k (SG-EMTF.cu:60, :462-486, :1601-1615). */
int emtf_get_elites(emtf_handle* h, int32_t k, float* fitness, float* x, int32_t x_stride);

/* emtf_get_elites without any device allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The ranked rows are staged in a fixed
 * array of 2^20 floats (4 MiB) that the library keeps in every GPU's memory until the
 * process ends: n_tasks * k * (x_stride + 1) floats at most (x_stride counts 0 when x is
 * NULL); larger requests return EMTF_E_INVALID.  Every call on one GPU shares that array
 * and overwrites it, so calls that overlap on one GPU (from several threads or handles)
 * race.  Same results as emtf_get_elites; synchronises. */
int emtf_get_elites_noalloc(emtf_handle* h, int32_t k, float* fitness, float* x, int32_t x_stride);

/* Convergence trace recorded on the device (the reference's per-task record, SG-EMTF_DE.cu:
 * 1793-1811: best at generation g0 and after every `interval` generations, then the worst
 * of the population at g0).  Records point 0 now (g0 = the handle's generation) and then,
 * during emtf_run, the best of every task at g0 + k interval, up to max_points points; each
 * point is one small kernel after the epoch that ends there, no host synchronisation.
 * interval 0 stops recording (the points stay readable); a new start replaces the trace. */
int emtf_trace_start(emtf_handle* h, int32_t interval, int32_t max_points, void* stream);

/* Points recorded so far: *n_points, best[n_points][n_tasks] and initial_worst[n_tasks]
 * (either may be NULL); synchronises.  Best / worst ignore NaN fitness. */
int emtf_trace(emtf_handle* h, int32_t* n_points, float* best, float* initial_worst);

int emtf_destroy(emtf_handle* h);

/* The engine's random draws (device pointers, asynchronous on `stream`): out[i][0..3] =
 * the Philox4x32-10 block that the engine draws at address addrs[i] = {task, generation,
 * purpose, index, sub, block} under `seed` (key = seed halves; counter = {block,
 * purpose<<24 | (sub & 0xFF)<<16 | (index & 0xFFFF), generation, task}).  out is 16-byte
 * aligned.  Lets a checker consume the exact streams of a run.  Replaces the per-thread
 * cuRAND XORWOW states (SG-EMTF.cu:163-179), which had no export. */
int emtf_eval_draws(uint64_t seed, int32_t n, const uint32_t* addrs, uint32_t* out, void* stream);

/* ---- Objective and adjoint (device pointers, asynchronous on `stream`) -------
 * f[i] = f(Q (x_i - o)) for n points x[n][dim], dim 1 .. 256, bit for bit the engine's
 * evaluator (rot NULL: identity; shift NULL: zero).  Replaces the reference's evaluation()
 * (proposed model/SG-EMTF_DE.cu:834-1069) as a reusable operator. */
int emtf_objective_fwd(int32_t func, int32_t dim, int32_t n, const float* x, const float* rot, const float* shift,
                       float* f, void* stream);
/* gx[i] = gf[i] * Q^T (df/dz)(Q (x_i - o)): the hand-written adjoint in binary32, dim 1 .. 128.
 * No counterpart in the reference (it has no gradients); used by the gradient polish
 * and by the torch.autograd.Function wrapper. */
int emtf_objective_adjoint(int32_t func, int32_t dim, int32_t n, const float* x, const float* rot,
                           const float* shift, const float* gf, float* gx, void* stream);

/* ---- CVRP genetic algorithm (exact integers) -----------------------------
 * Replaces the CVRP variant of the engine (proposed model/SG-EMTF_cvrp.cu: its
 * pop_init :155-187, para_init :210-225, CrossOver :668-738, evaluation
 * :859-962 and iter() :1495-1500).  Genome: a giant tour (uint8 customer ids
 * 1..n); decode: Prins split; operators: OX crossover, swap / 2-opt mutation;
 * distances nint(sqrt(dx^2+dy^2)) of integer coordinates. */
typedef struct emtf_cvrp_instance {
    int32_t n;               /* customers, 2..255 (depot is node 0) */
    int32_t capacity;        /* vehicle capacity, >= every demand, < 64 x smallest demand */
    int32_t pop;             /* individuals, <= 256, a multiple of n_islands */
    const int32_t* xy;       /* (n+1) x 2 integer coordinates, row 0 = depot */
    const int32_t* demand;   /* n+1 demands, demand[0] = 0, others in [1, 255] */
} emtf_cvrp_instance;

typedef struct emtf_cvrp_params {
    float p_swap, p_2opt;    /* per child: swap mutation with p_swap, else 2-opt with p_2opt */
    int32_t n_islands, migrate_interval, migrate_count;
    uint64_t seed;
} emtf_cvrp_params;

typedef struct emtf_cvrp_handle emtf_cvrp_handle;

int emtf_cvrp_create(const emtf_cvrp_instance* instances, int32_t n_instances, const emtf_cvrp_params* params,
                     emtf_cvrp_handle** out);
int emtf_cvrp_run(emtf_cvrp_handle* h, int32_t generations, void* stream);
int emtf_cvrp_sync(emtf_cvrp_handle* h);
/* tours[pop][n] and integer costs[pop] of one instance (synchronises) */
int emtf_cvrp_get_population(emtf_cvrp_handle* h, int32_t instance, uint8_t* tours, int32_t* costs);
/* best cost of every instance and its tour (rows of `stride` >= n) */
int emtf_cvrp_get_best(emtf_cvrp_handle* h, int32_t* costs, uint8_t* tours, int32_t stride);
int emtf_cvrp_destroy(emtf_cvrp_handle* h);

/* Message of the calling thread's last failure ("" if none). */
const char* emtf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
