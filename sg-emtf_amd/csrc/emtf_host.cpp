// emtf_host.cpp -- host runtime behind include/emtf.h.
//
// Owns the device state of a batch of tasks (one contiguous arena per array,
// per-task offsets in a descriptor table) and schedules epochs of the island
// kernels.  Tasks are grouped into buckets by the kernel that runs them: the
// LDS-resident kernel built for the smallest row geometry KB >= ceil(dim / 8)
// whose population fits LDS, or else the wide kernel (rows in global memory)
// for the smallest 32-row tile count NT >= dim / 32.  Each epoch launches every
// bucket on its own stream, forked from and joined back into the caller's
// stream.  An epoch never crosses a transfer boundary, so the epoch boundary is
// the synchronisation point of the elite mailbox: the launches that end at
// generation g write mailbox buffer (g / interval) & 1, the next ones read it
// while writing the other one.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/emtf.h"
#include "emtf_types.h"

// One launcher per (gene blocks per parity half KB, algorithm family); island.hip is
// built once per pair.  A batch uses the smallest KB >= ceil(max dim / 8).
#define EMTF_DECL_LAUNCH(kb, al)                                                                          \
    extern "C" int emtf_launch_epoch_kb##kb##_a##al(const emtf::EpochArgs* args, const int* order, int n, \
                                                    int lds_bytes, hipStream_t stream);
#define EMTF_DECL_PAIR(kb) EMTF_DECL_LAUNCH(kb, 0) EMTF_DECL_LAUNCH(kb, 2)
EMTF_DECL_PAIR(2) EMTF_DECL_PAIR(4) EMTF_DECL_PAIR(7) EMTF_DECL_PAIR(8) EMTF_DECL_PAIR(12) EMTF_DECL_PAIR(13)
EMTF_DECL_PAIR(16)
extern "C" int emtf_lds_floats_kb(int pop, int kb);
#define EMTF_DECL_WIDE(nt, al)                                                                             \
    extern "C" int emtf_launch_wide_nt##nt##_a##al(const emtf::EpochArgs* args, const int* order, int n,   \
                                                   int lds_bytes, hipStream_t stream);                     \
    extern "C" int emtf_launch_wide8_nt##nt##_a##al(const emtf::EpochArgs* args, const int* order, int n,  \
                                                    int lds_bytes, hipStream_t stream);                    \
    extern "C" int emtf_wide_lds_floats_nt##nt##_a##al(int pop);
#define EMTF_DECL_WPAIR(nt) EMTF_DECL_WIDE(nt, 0) EMTF_DECL_WIDE(nt, 2)
EMTF_DECL_WPAIR(2) EMTF_DECL_WPAIR(4) EMTF_DECL_WPAIR(5) EMTF_DECL_WPAIR(6) EMTF_DECL_WPAIR(7) EMTF_DECL_WPAIR(8)
// the wide kernel with the task's Q image in LDS (NT <= 6, 512 threads)
#define EMTF_DECL_WIDEQ(nt, al)                                                                            \
    extern "C" int emtf_launch_wideq_nt##nt##_a##al(const emtf::EpochArgs* args, const int* order, int n,  \
                                                    int lds_bytes, hipStream_t stream);                    \
    extern "C" int emtf_wideq_lds_floats_nt##nt##_a##al(int pop);
#define EMTF_DECL_QPAIR(nt) EMTF_DECL_WIDEQ(nt, 0) EMTF_DECL_WIDEQ(nt, 2)
EMTF_DECL_QPAIR(2) EMTF_DECL_QPAIR(4) EMTF_DECL_QPAIR(5) EMTF_DECL_QPAIR(6)
#define EMTF_DECL_OBJ(kb)                                                                                   \
    extern "C" int emtf_launch_objective_kb##kb(const emtf::ObjArgs* a, int adjoint, hipStream_t s);      \
    extern "C" int emtf_polish_lds_floats_kb##kb(int pop, int waves, int cpe);                             \
    extern "C" int emtf_launch_polish_kb##kb(const emtf::EpochArgs* args, const int* order, int n, int lds_bytes, \
                                             hipStream_t stream);
EMTF_DECL_OBJ(2) EMTF_DECL_OBJ(4) EMTF_DECL_OBJ(7) EMTF_DECL_OBJ(8) EMTF_DECL_OBJ(12) EMTF_DECL_OBJ(13) EMTF_DECL_OBJ(16)
// forward objective for 128 < dim <= 32 NT (Q and points read from global memory)
#define EMTF_DECL_OBJW(nt) extern "C" int emtf_launch_objective_nt##nt##_fwd(const emtf::ObjArgs* a, hipStream_t s);
EMTF_DECL_OBJW(5) EMTF_DECL_OBJW(6) EMTF_DECL_OBJW(7) EMTF_DECL_OBJW(8)

// the persistent schedule's kernel (LDS kernel builds) and its occupancy query
#define EMTF_DECL_PERSIST(kb, al)                                                                           \
    extern "C" int emtf_launch_persist_kb##kb##_a##al(const emtf::EpochArgs* args, const emtf::PersistArgs* pa, \
                                                      const int* order, int blocks, int lds_bytes,         \
                                                      hipStream_t stream);
#define EMTF_DECL_PPAIR(kb) EMTF_DECL_PERSIST(kb, 0) EMTF_DECL_PERSIST(kb, 2)
EMTF_DECL_PPAIR(2) EMTF_DECL_PPAIR(4) EMTF_DECL_PPAIR(7) EMTF_DECL_PPAIR(8) EMTF_DECL_PPAIR(12) EMTF_DECL_PPAIR(13)
EMTF_DECL_PPAIR(16)
typedef int (*persist_fn)(const emtf::EpochArgs*, const emtf::PersistArgs*, const int*, int, int, hipStream_t);
extern "C" int emtf_launch_fill_i32(int32_t* p, int n, int32_t value, hipStream_t stream);

typedef int (*launch_fn)(const emtf::EpochArgs*, const int*, int, int, hipStream_t);
extern "C" int emtf_launch_trace(const emtf::TaskDev* tasks, int n_tasks, const float* fit, float* out_best,
                                 float* out_worst, hipStream_t stream);
extern "C" int emtf_launch_draws(uint64_t seed, int n, const uint32_t* addr, uint32_t* out, hipStream_t stream);
extern "C" int emtf_launch_elites(const emtf::TaskDev* tasks, int n_tasks, const float* X, const float* fit, int k,
                                  int stride, float* out_x, float* out_f, hipStream_t stream);
extern "C" int emtf_scratch_resolve(int device, float** out, size_t* capacity);
extern "C" int emtf_launch_gen_counter(int32_t* gen, int32_t value, int32_t add, hipStream_t stream);
extern "C" int emtf_launch_donors(const emtf::TaskDev* tasks, int n_tasks, int g, int32_t* list, hipStream_t stream);

static const int kKB[] = {2, 4, 7, 8, 12, 13, 16};

static const int kNT[] = {2, 4, 5, 6, 7, 8};

static int pick_kb(int max_dim) {
    const int need = (max_dim + 7) / 8;
    for (int kb : kKB)
        if (kb >= need) return kb;
    return -1;
}

static int pick_nt(int dim) {
    for (int nt : kNT)
        if (32 * nt >= dim) return nt;
    return -1;
}

// the 512-thread build of the wide kernel (populations of more than four tiles)
static launch_fn pick_wide8(int nt, bool ga) {
#define EMTF_CASE(k) \
    case k: return ga ? emtf_launch_wide8_nt##k##_a2 : emtf_launch_wide8_nt##k##_a0;
    switch (nt) {
        EMTF_CASE(2) EMTF_CASE(4) EMTF_CASE(5) EMTF_CASE(6) EMTF_CASE(7) EMTF_CASE(8)
    }
#undef EMTF_CASE
    return nullptr;
}

static launch_fn pick_wideq(int nt, bool ga, int (**lds)(int)) {
#define EMTF_CASE(k)                                                                 \
    case k:                                                                          \
        *lds = ga ? emtf_wideq_lds_floats_nt##k##_a2 : emtf_wideq_lds_floats_nt##k##_a0; \
        return ga ? emtf_launch_wideq_nt##k##_a2 : emtf_launch_wideq_nt##k##_a0;
    switch (nt) {
        EMTF_CASE(2) EMTF_CASE(4) EMTF_CASE(5) EMTF_CASE(6)
    }
#undef EMTF_CASE
    return nullptr;
}

static launch_fn pick_wide(int nt, bool ga, int (**lds)(int)) {
#define EMTF_CASE(k)                                                               \
    case k:                                                                        \
        *lds = ga ? emtf_wide_lds_floats_nt##k##_a2 : emtf_wide_lds_floats_nt##k##_a0; \
        return ga ? emtf_launch_wide_nt##k##_a2 : emtf_launch_wide_nt##k##_a0;
    switch (nt) {
        EMTF_CASE(2) EMTF_CASE(4) EMTF_CASE(5) EMTF_CASE(6) EMTF_CASE(7) EMTF_CASE(8)
    }
#undef EMTF_CASE
    return nullptr;
}

static launch_fn pick_launcher(int kb, bool ga) {
#define EMTF_CASE(k) \
    case k: return ga ? emtf_launch_epoch_kb##k##_a2 : emtf_launch_epoch_kb##k##_a0;
    switch (kb) {
        EMTF_CASE(2) EMTF_CASE(4) EMTF_CASE(7) EMTF_CASE(8) EMTF_CASE(12) EMTF_CASE(13) EMTF_CASE(16)
    }
#undef EMTF_CASE
    return nullptr;
}

static persist_fn pick_persist(int kb, bool ga) {
#define EMTF_CASE(k) \
    case k: return ga ? emtf_launch_persist_kb##k##_a2 : emtf_launch_persist_kb##k##_a0;
    switch (kb) {
        EMTF_CASE(2) EMTF_CASE(4) EMTF_CASE(7) EMTF_CASE(8) EMTF_CASE(12) EMTF_CASE(13) EMTF_CASE(16)
    }
#undef EMTF_CASE
    return nullptr;
}

// Epochs one persistent launch holds (its per-epoch read counters), and the bound of one
// device-side wait (10 s of the 100 MHz wall clock): an item waits at most for another
// item's epoch, far below it.
constexpr int kPersistEpochs = 1024;
constexpr int64_t kPersistWaitTicks = 1000000000LL;

typedef int (*polish_lds_fn)(int, int, int);

static launch_fn pick_polish(int kb, polish_lds_fn* lds) {
#define EMTF_CASE(k)                          \
    case k:                                   \
        *lds = emtf_polish_lds_floats_kb##k; \
        return emtf_launch_polish_kb##k;
    switch (kb) {
        EMTF_CASE(2) EMTF_CASE(4) EMTF_CASE(7) EMTF_CASE(8) EMTF_CASE(12) EMTF_CASE(13) EMTF_CASE(16)
    }
#undef EMTF_CASE
    return nullptr;
}

// Relative cost of one generation per (individual, dimension), used to
// dispatch the expensive tasks first.
static double func_cost(int f) {
    switch (f) {
    case EMTF_WEIERSTRASS: case emtf::F_REF_WEIERSTRASS: return 21.0;
    case EMTF_SCHWEFEL: case EMTF_ACKLEY: case EMTF_RASTRIGIN: case EMTF_GRIEWANK: return 2.0;
    default: return 1.0;
    }
}

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_OR_FAIL(expr)                                                                            \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return fail(EMTF_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int kMaxLdsBytes = 160 * 1024;

}  // namespace

// Tasks run by one kernel variant: d_order[offset .. offset + count), most expensive first.
struct Variant {
    bool wide = false;
    int geo = 0;                 // KB (LDS kernel) or NT (wide kernel)
    launch_fn fn = nullptr;
    int lds_bytes = 0;
    const char* name = "";
};

constexpr int kVariants = 3;

struct Bucket {
    bool wide = false;
    int geo = 0;                 // KB (LDS path) or NT (wide path)
    launch_fn fn = nullptr;
    int lds_bytes = 0;
    int offset = 0, count = 0;
    double cost = 0.0;
    int polish_lds_bytes = 0;
    int polish_waves = 0;        // 0: polish off
    int polish_cpe = 1;
    int stream = 0;              // index into the handle's streams (0 = the caller's)
    Variant v[kVariants];        // [0] primary, [1] [2] alternatives (fn null if none)
    int use = 0;                 // variant launched
    float ms[kVariants] = {};    // calibration times of the variants
};

struct emtf_handle {
    int T = 0;
    emtf::EngineParams p{};
    std::vector<emtf::TaskDev> tasks;
    emtf::TaskDev* d_tasks = nullptr;
    float* d_X = nullptr;
    float* d_fit = nullptr;
    float* d_Q = nullptr;
    float* d_shift = nullptr;
    float* d_tab = nullptr;
    float* d_mailbox = nullptr;
    float* d_scratch = nullptr;
    float* d_qw = nullptr;       // padded Q images of the wide tasks
    int* d_order = nullptr;      // task ids grouped by bucket
    std::vector<Bucket> buckets;
    bool polish = false;
    int64_t mailbox_stride = 0;
    int gen = 0;
    int max_dim = 0;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> side;  // extra streams for buckets 1.. (non-blocking, default priority)
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;
    hipEvent_t ev_order = nullptr;  // orders a run on a new stream after the previous one
    // convergence trace: best of every task at trace_g0 + k trace_interval, k < trace_max
    int trace_interval = 0, trace_max = 0, trace_points = 0, trace_g0 = 0;
    float* d_trace = nullptr;        // [trace_max][T]
    float* d_trace_worst = nullptr;  // [T], the worst fitness at trace_g0
    // elite export staging, kept between calls (grown on demand)
    float* d_elite = nullptr;
    size_t elite_cap = 0;        // floats
    int device = 0;              // the GPU the handle lives on
    // replayed epochs: one epoch's launches (fork, every launch unit, join, counter advance)
    // captured once into a hipGraph and replayed while the epochs keep that length
    bool use_graph = false;
    int32_t* d_gen = nullptr;    // device generation counter the replayed kernels start from
    bool gen_valid = false;      // d_gen == gen (false after directly launched epochs)
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_len = 0;           // generations per replayed epoch
    // online admission: tasks start at emtf_admit; the generation counter is the clock
    bool online = false;
    std::vector<int> ends;       // clock values at which admitted tasks finish (sorted)
    int32_t* d_donors = nullptr; // [T] donor list of a transfer boundary, then its count
    // persistent schedule: its launcher and the device counters (error flag, item counter,
    // T stamps, kPersistEpochs read counts)
    persist_fn persist = nullptr;
    int32_t* d_persist = nullptr;
};

static void free_handle(emtf_handle* h) {
    if (!h) return;
    (void)hipFree(h->d_tasks);
    (void)hipFree(h->d_X);
    (void)hipFree(h->d_fit);
    (void)hipFree(h->d_Q);
    (void)hipFree(h->d_shift);
    (void)hipFree(h->d_tab);
    (void)hipFree(h->d_mailbox);
    (void)hipFree(h->d_scratch);
    (void)hipFree(h->d_qw);
    (void)hipFree(h->d_order);
    (void)hipFree(h->d_trace);
    (void)hipFree(h->d_trace_worst);
    (void)hipFree(h->d_elite);
    (void)hipFree(h->d_gen);
    (void)hipFree(h->d_donors);
    (void)hipFree(h->d_persist);
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
    for (hipStream_t s : h->side) (void)hipStreamDestroy(s);

    for (hipEvent_t e : h->ev_join) (void)hipEventDestroy(e);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_order) (void)hipEventDestroy(h->ev_order);
    delete h;
}

// Work queued on a stream other than the previous call's waits for that call's work: a
// handle's launches stay in call order whichever streams the caller passes.
static int use_stream(emtf_handle* h, hipStream_t s) {
    if (s != h->stream) {
        if (!h->ev_order) HIP_OR_FAIL(hipEventCreateWithFlags(&h->ev_order, hipEventDisableTiming));
        HIP_OR_FAIL(hipEventRecord(h->ev_order, h->stream));
        HIP_OR_FAIL(hipStreamWaitEvent(s, h->ev_order, 0));
        h->stream = s;
    }
    return EMTF_OK;
}

// The reference's Weierstrass constant Weierstrass_para (SG-EMTF.cu:354-361): sum of
// 0.5^k cos(2 PI 3^k 0.5) over k = 0..20 in binary32, PI = 3.1415926, the argument formed
// in double and rounded to binary32, its cosine correctly rounded to binary32.
static float weierstrass_para_ref() {
    static const float p3[21] = {1.f, 3.f, 9.f, 27.f, 81.f, 243.f, 729.f, 2187.f, 6561.f, 19683.f,
                                 59049.f, 177147.f, 531441.f, 1594323.f, 4782969.f, 14348907.f,
                                 43046721.f, 129140163.f, 387420489.f, 1162261467.f, 3486784401.f};
    float a = 1.0f, res = 0.0f;
    for (int k = 0; k <= 20; ++k) {
        const double arg = 2.0 * 3.1415926 * (double)p3[k] * 0.5;
        const float c = (float)std::cos((double)(float)arg);
        const float t = a * c;
        res = res + t;
        a = a * 0.5f;
    }
    return res;
}

// Per-dimension table: Elliptic weights 10^(6 j/(D-1)), Griewank 1/(2 pi sqrt(j+1)), the
// Schwefel constant (textbook 418.982887, reference 418.9829); reference forms: Griewank
// divisors sqrtf(j+1) (+inf for the GA copy, whose product is disabled) and the
// Weierstrass constant.
static void func_table(int func, int dim, float* t) {
    const float w_ref = func == emtf::F_REF_WEIERSTRASS ? weierstrass_para_ref() : 0.0f;
    for (int j = 0; j < dim; ++j) {
        double v = 0.0;
        if (func == EMTF_ELLIPTIC) v = (dim > 1) ? std::pow(1.0e6, (double)j / (double)(dim - 1)) : 1.0;
        else if (func == EMTF_GRIEWANK) v = 1.0 / (2.0 * M_PI * std::sqrt((double)(j + 1)));
        t[j] = (float)v;
        if (func == EMTF_SCHWEFEL) t[j] = 418.982887f;
        else if (func == emtf::F_REF_SCHWEFEL) t[j] = 418.9829f;
        else if (func == emtf::F_REF_GRIEWANK) t[j] = std::sqrt((float)(j + 1));
        else if (func == emtf::F_REF_GRIEWANK_GA) t[j] = INFINITY;
        else if (func == emtf::F_REF_WEIERSTRASS) t[j] = w_ref;
    }
}

// The objective id the kernels run for a public id under the chosen formulas.
static int kernel_func(int func, int formulas) {
    if (formulas == EMTF_FORMULAS_TEXTBOOK) return func;
    switch (func) {
    case EMTF_ACKLEY: return emtf::F_REF_ACKLEY;
    case EMTF_RASTRIGIN: return emtf::F_REF_RASTRIGIN;
    case EMTF_GRIEWANK: return formulas == EMTF_FORMULAS_REFERENCE_GA ? emtf::F_REF_GRIEWANK_GA : emtf::F_REF_GRIEWANK;
    case EMTF_WEIERSTRASS: return emtf::F_REF_WEIERSTRASS;
    case EMTF_SCHWEFEL: return emtf::F_REF_SCHWEFEL;
    default: return func;
    }
}

static bool polish_at(const emtf_handle* h, int g) {
    return h->polish && g > 0 && g % h->p.polish_interval == 0;
}

static bool publish_at(const emtf_handle* h, int g) {
    const int TI = h->p.transfer_interval;
    return TI > 0 && h->T > 1 && h->p.transfer_count > 0 && g > 0 && g % TI == 0;
}

// Epoch g0 -> g1 (or the initial evaluation) of every bucket, each on its stream;
// the polish of generation g1 follows in its own kernel, which then also
// publishes the elites.
static int launch(emtf_handle* h, int g0, int g1, int init, hipStream_t stream, const int32_t* gen_dev = nullptr) {
    emtf::EpochArgs a{};
    a.gen_dev = gen_dev;
    a.tasks = h->d_tasks;
    a.n_tasks = h->T;
    a.g0 = g0;
    a.g1 = g1;
    a.init = init;
    a.X = h->d_X;
    a.fit = h->d_fit;
    a.Q = h->d_Q;
    a.shift = h->d_shift;
    a.tab = h->d_tab;
    a.mailbox = h->d_mailbox;
    a.mailbox_stride = h->mailbox_stride;
    a.scratch = h->d_scratch;
    a.Qw = h->d_qw;
    a.p = h->p;
    if (h->online) {
        // the donors of this boundary, listed on the device before the epoch's kernels
        a.donors = h->d_donors;
        a.n_donors = h->d_donors + h->T;
        const int TI = h->p.transfer_interval;
        if (!init && TI > 0 && h->p.transfer_count > 0 && g0 > 0 && g0 % TI == 0) {
            const int e = emtf_launch_donors(h->d_tasks, h->T, g0, h->d_donors, stream);
            if (e != 0) return fail(EMTF_E_DEVICE, std::string("donor list: ") + hipGetErrorString((hipError_t)e));
        }
    }
    const bool pol = !init && polish_at(h, g1);
    const bool pub = !init && publish_at(h, g1);
    const bool fork = !h->side.empty();
    if (fork) {
        HIP_OR_FAIL(hipEventRecord(h->ev_fork, stream));
        for (hipStream_t s : h->side) HIP_OR_FAIL(hipStreamWaitEvent(s, h->ev_fork, 0));
    }
    for (Bucket& b : h->buckets) {
        hipStream_t s = b.stream == 0 ? stream : h->side[b.stream - 1];
        a.publish = (pub && !(pol && b.polish_waves > 0)) ? 1 : 0;
        const Variant& var = b.v[b.use];
        int rc = var.fn(&a, h->d_order + b.offset, b.count, var.lds_bytes, s);
        if (rc != 0) return fail(EMTF_E_DEVICE, std::string("island kernel launch: ") + hipGetErrorString((hipError_t)rc));
        if (pol && b.polish_waves > 0) {
            int (*plds)(int, int, int) = nullptr;
            launch_fn pf = pick_polish(b.geo, &plds);
            emtf::EpochArgs ap = a;
            ap.publish = pub ? 1 : 0;
            ap.polish_waves = b.polish_waves;
            ap.polish_cpe = b.polish_cpe;
            rc = pf(&ap, h->d_order + b.offset, b.count, b.polish_lds_bytes, s);
            if (rc != 0) return fail(EMTF_E_DEVICE, std::string("polish kernel launch: ") + hipGetErrorString((hipError_t)rc));
        }
    }
    if (fork) {
        for (size_t k = 0; k < h->side.size(); ++k) {
            HIP_OR_FAIL(hipEventRecord(h->ev_join[k], h->side[k]));
            HIP_OR_FAIL(hipStreamWaitEvent(stream, h->ev_join[k], 0));
        }
    }
    return EMTF_OK;
}

// Capture one epoch of `len` generations (its first at g0) into the handle's graph: the
// fork, every launch unit and the join as launch() issues them, with the kernels starting
// from the device counter, then the counter's advance by len.  Replaces a previous graph.
static int capture_epoch(emtf_handle* h, int g0, int len) {
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->graph_len = 0;
    if (!h->cap_stream) HIP_OR_FAIL(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
    HIP_OR_FAIL(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeRelaxed));
    int rc = launch(h, g0, g0 + len, 0, h->cap_stream, h->d_gen);
    if (rc == EMTF_OK) {
        const int e = emtf_launch_gen_counter(h->d_gen, len, 1, h->cap_stream);
        if (e != 0) rc = fail(EMTF_E_DEVICE, std::string("counter: ") + hipGetErrorString((hipError_t)e));
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(h->cap_stream, &g);
    if (rc != EMTF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return fail(EMTF_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    h->graph = g;
    HIP_OR_FAIL(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
    h->graph_len = len;
    return EMTF_OK;
}

// Epoch g0 -> g0 + len replayed from the graph (captured on first use or when len changes).
static int replay_epoch(emtf_handle* h, int g0, int len, hipStream_t s) {
    if (!h->graph_exec || h->graph_len != len)
        if (const int rc = capture_epoch(h, g0, len); rc != EMTF_OK) return rc;
    if (!h->gen_valid) {
        const int e = emtf_launch_gen_counter(h->d_gen, g0, 0, s);
        if (e != 0) return fail(EMTF_E_DEVICE, std::string("counter: ") + hipGetErrorString((hipError_t)e));
    }
    HIP_OR_FAIL(hipGraphLaunch(h->graph_exec, s));
    h->gen_valid = true;
    return EMTF_OK;
}

// Autotuner decisions kept for later creates of batches with the same launch units on the
// same kind of device: in memory for the life of the process and, when the environment
// variable EMTF_TUNE_CACHE names a file, in that file (one line per batch signature,
// appended), so that another process skips the calibration too.
namespace {
std::mutex g_tune_mu;
std::map<std::string, std::string> g_tune;  // batch signature -> per unit "use ms0 ms1 ms2"
bool g_tune_file_read = false;

const char* tune_file() {
    const char* f = getenv("EMTF_TUNE_CACHE");
    return f && *f ? f : nullptr;
}
}  // namespace

// What the calibration of a batch depends on: the device kind, the knobs that change a
// generation's work, and per launch unit its variants, task count (rounded up to a power
// of two), largest dimension and population and objective-cost sum.
static std::string tune_signature(const emtf_handle* h, const std::vector<emtf::TaskDev>& tasks,
                                  const std::vector<int>& order) {
    hipDeviceProp_t prop{};
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipGetDeviceProperties(&prop, dev);
    char buf[256];
    snprintf(buf, sizeof buf, "emtf-tune-1 %s/%d a%d i%d x%d m%d t%d", prop.gcnArchName, prop.multiProcessorCount,
             h->p.algo, h->p.n_islands, h->p.exchange_interval, h->p.migrate_interval, h->p.transfer_interval);
    std::string sig = buf;
    for (const Bucket& b : h->buckets) {
        int maxd = 0, maxp = 0, n2 = 1;
        double c = 0.0;
        for (int i = b.offset; i < b.offset + b.count; ++i) {
            const emtf::TaskDev& td = tasks[order[i]];
            maxd = std::max(maxd, td.dim);
            maxp = std::max(maxp, td.pop);
            c += func_cost(td.func);
        }
        while (n2 < b.count) n2 *= 2;
        snprintf(buf, sizeof buf, " |%d.%d n%d d%d p%d c%.0f", b.wide ? 1 : 0, b.geo, n2, maxd, maxp, c * 64.0 / b.count);
        sig += buf;
        for (int v = 0; v < kVariants; ++v)
            if (b.v[v].fn) {
                snprintf(buf, sizeof buf, " v%d:%d.%d.%d", v, b.v[v].wide ? 1 : 0, b.v[v].geo, b.v[v].lds_bytes);
                sig += buf;
            }
    }
    return sig;
}

static bool tune_lookup(emtf_handle* h, const std::string& sig) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (!g_tune_file_read) {
        g_tune_file_read = true;
        if (const char* f = tune_file())
            if (FILE* fp = fopen(f, "r")) {
                char line[8192];
                while (fgets(line, sizeof line, fp)) {
                    char* tab = strchr(line, '\t');
                    if (!tab) continue;
                    *tab = 0;
                    std::string val = tab + 1;
                    while (!val.empty() && (val.back() == '\n' || val.back() == '\r')) val.pop_back();
                    g_tune[line] = val;
                }
                fclose(fp);
            }
    }
    auto it = g_tune.find(sig);
    if (it == g_tune.end()) return false;
    // parse first, apply only a complete, consistent record
    std::vector<int> use;
    std::vector<float> ms;
    const char* q = it->second.c_str();
    for (size_t k = 0; k < h->buckets.size(); ++k) {
        int u = 0, n = 0;
        float m[kVariants] = {};
        if (sscanf(q, " %d %f %f %f%n", &u, &m[0], &m[1], &m[2], &n) != 4) return false;
        if (u < 0 || u >= kVariants || !h->buckets[k].v[u].fn) return false;
        use.push_back(u);
        for (int v = 0; v < kVariants; ++v) ms.push_back(m[v]);
        q += n;
    }
    for (size_t k = 0; k < h->buckets.size(); ++k) {
        h->buckets[k].use = use[k];
        for (int v = 0; v < kVariants; ++v) h->buckets[k].ms[v] = ms[k * kVariants + v];
    }
    return true;
}

static void tune_store(const emtf_handle* h, const std::string& sig) {
    std::string val;
    char buf[96];
    for (const Bucket& b : h->buckets) {
        snprintf(buf, sizeof buf, " %d %.6g %.6g %.6g", b.use, b.ms[0], b.ms[1], b.ms[2]);
        val += buf;
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[sig] = val;
    if (const char* f = tune_file())
        if (FILE* fp = fopen(f, "a")) {
            fprintf(fp, "%s\t%s\n", sig.c_str(), val.c_str());
            fclose(fp);
        }
}

// Units to streams longest-first on the measured times (the caller's stream and up to
// three more), and the tuner's log line (EMTF_TUNE_LOG=1).
static void assign_streams(emtf_handle* h, bool cached) {
    const int n_streams = 1 + (int)h->side.size();
    std::stable_sort(h->buckets.begin(), h->buckets.end(),
                     [](const Bucket& x, const Bucket& y) { return x.ms[x.use] > y.ms[y.use]; });
    std::vector<double> load(n_streams, 0.0);
    for (Bucket& b : h->buckets) {
        int s = 0;
        for (int q = 1; q < n_streams; ++q)
            if (load[q] < load[s]) s = q;
        b.stream = s;
        load[s] += b.ms[b.use];
    }
    if (const char* lg = getenv("EMTF_TUNE_LOG"); lg && *lg == '1')
        for (const Bucket& b : h->buckets) {
            fprintf(stderr, "emtf tune: %d tasks on stream %d:", b.count, b.stream);
            for (int v = 0; v < kVariants; ++v)
                if (b.v[v].fn) fprintf(stderr, " [%d] %s %d %.3f ms", v, b.v[v].name, b.v[v].geo, b.ms[v]);
            fprintf(stderr, " -> [%d]%s\n", b.use, cached ? " (cached)" : "");
        }
}

// Autotuner: with the batch initialised, time every launch unit alone for a few
// generations with each of its kernel variants (the LDS kernel and the wide kernel
// compute the same bits), keep the faster, and give the units to the streams
// longest-first on the measured times.  Populations and fitness are restored after
// each trial (the trials start at generation 0 and publish nothing), so the run's
// results do not depend on the calibration.  A batch whose signature was calibrated
// before (this process, or the EMTF_TUNE_CACHE file) takes the recorded decisions.
static int calibrate(emtf_handle* h, int64_t nX, int64_t nF, const std::vector<int>& order) {
    bool any_alt = false;
    for (const Bucket& b : h->buckets)
        for (int v = 1; v < kVariants; ++v) any_alt = any_alt || b.v[v].fn;
    if (!any_alt && h->buckets.size() < 2) return EMTF_OK;
    const std::string sig = tune_signature(h, h->tasks, order);
    if (tune_lookup(h, sig)) {
        assign_streams(h, true);
        return EMTF_OK;
    }
    const int gens = 4;
    float *bx = nullptr, *bf = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void**)&bx, sizeof(float) * (size_t)std::max<int64_t>(nX, 1));
    if (e == hipSuccess) e = hipMalloc((void**)&bf, sizeof(float) * (size_t)std::max<int64_t>(nF, 1));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemcpy(bx, h->d_X, sizeof(float) * (size_t)nX, hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(bf, h->d_fit, sizeof(float) * (size_t)nF, hipMemcpyDeviceToDevice);
    emtf::EpochArgs a{};
    a.tasks = h->d_tasks;
    a.n_tasks = h->T;
    a.g0 = 0;
    a.g1 = gens;
    a.X = h->d_X;
    a.fit = h->d_fit;
    a.Q = h->d_Q;
    a.shift = h->d_shift;
    a.tab = h->d_tab;
    a.mailbox = h->d_mailbox;
    a.mailbox_stride = h->mailbox_stride;
    a.scratch = h->d_scratch;
    a.Qw = h->d_qw;
    a.p = h->p;
    a.publish = 0;
    for (Bucket& b : h->buckets) {
        // two trials per variant, the faster kept: a kernel's first launch in the process
        // also pays for loading its code object
        for (int v = 0; v < kVariants && e == hipSuccess; ++v) {
            if (!b.v[v].fn) continue;
            for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
                float ms = 0.0f;
                e = hipEventRecord(e0, nullptr);
                if (e == hipSuccess)
                    e = (hipError_t)b.v[v].fn(&a, h->d_order + b.offset, b.count, b.v[v].lds_bytes, nullptr);
                if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e == hipSuccess) e = hipMemcpy(h->d_X, bx, sizeof(float) * (size_t)nX, hipMemcpyDeviceToDevice);
                if (e == hipSuccess) e = hipMemcpy(h->d_fit, bf, sizeof(float) * (size_t)nF, hipMemcpyDeviceToDevice);
                b.ms[v] = rep == 0 ? ms : std::min(b.ms[v], ms);
            }
        }
        b.use = 0;
        for (int v = 1; v < kVariants; ++v)
            if (b.v[v].fn && b.ms[v] > 0.0f && b.ms[v] < b.ms[b.use]) b.use = v;
    }
    (void)hipFree(bx);
    (void)hipFree(bf);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(EMTF_E_DEVICE, std::string("calibration: ") + hipGetErrorString(e));
    tune_store(h, sig);
    assign_streams(h, false);
    return EMTF_OK;
}

extern "C" {

const char* emtf_last_error(void) { return g_err.c_str(); }

// shared with the CVRP runtime (cvrp_host.cpp); not part of the public header
void emtf_set_error(const char* msg) { g_err = msg ? msg : ""; }

int emtf_create(const emtf_task_desc* tasks, int32_t n_tasks, const emtf_params* prm, emtf_handle** out) {
    g_err.clear();
    if (!out) return fail(EMTF_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!tasks || !prm || n_tasks < 1) return fail(EMTF_E_INVALID, "tasks/params missing or n_tasks < 1");
    if (prm->algo < EMTF_DE_RAND_1_BIN || prm->algo > EMTF_GA_SBX_PM) return fail(EMTF_E_INVALID, "unknown algo");
    const int nis = prm->n_islands < 1 ? 1 : prm->n_islands;
    if (prm->transfer_count < 0 || prm->migrate_count < 0 || prm->transfer_interval < 0 ||
        prm->migrate_interval < 0 || prm->exchange_interval < 0)
        return fail(EMTF_E_INVALID, "negative interval or count");
    if (!(prm->F == prm->F) || !(prm->CR == prm->CR)) return fail(EMTF_E_INVALID, "F/CR is NaN");
    if (prm->polish_interval < 0 || prm->polish_steps < 0 || prm->polish_topk < 0)
        return fail(EMTF_E_INVALID, "negative polish knob");
    if (prm->schedule < EMTF_SCHEDULE_AUTO || prm->schedule > EMTF_SCHEDULE_PERSISTENT)
        return fail(EMTF_E_INVALID, "unknown schedule");
    if (prm->tune < 0 || prm->tune > 3)
        return fail(EMTF_E_INVALID, "tune must be 0 (calibrate), 1 (off), 2 or 3 (one alternative)");
    if (prm->formulas < EMTF_FORMULAS_TEXTBOOK || prm->formulas > EMTF_FORMULAS_REFERENCE_GA)
        return fail(EMTF_E_INVALID, "unknown formulas");
    if (prm->formulas != EMTF_FORMULAS_TEXTBOOK && prm->polish_interval > 0 && prm->polish_steps > 0 &&
        prm->polish_topk > 0)
        return fail(EMTF_E_UNSUPPORTED, "gradient polish needs the textbook formulas");
    if (prm->admission < EMTF_ADMIT_AT_CREATE || prm->admission > EMTF_ADMIT_ONLINE)
        return fail(EMTF_E_INVALID, "unknown admission");
    const bool online = prm->admission == EMTF_ADMIT_ONLINE;
    if (online && prm->polish_interval > 0 && prm->polish_steps > 0 && prm->polish_topk > 0)
        return fail(EMTF_E_UNSUPPORTED, "gradient polish is not available with online admission");

    int max_dim = 0;
    for (int t = 0; t < n_tasks; ++t) {
        const emtf_task_desc& d = tasks[t];
        char buf[160];
        if (d.func < 0 || d.func > EMTF_ELLIPTIC) { snprintf(buf, sizeof buf, "task %d: unknown func %d", t, d.func); return fail(EMTF_E_INVALID, buf); }
        if (d.dim < 1) { snprintf(buf, sizeof buf, "task %d: dim %d < 1", t, d.dim); return fail(EMTF_E_INVALID, buf); }
        if (d.dim > 256) { snprintf(buf, sizeof buf, "task %d: dim %d > 256 not built", t, d.dim); return fail(EMTF_E_UNSUPPORTED, buf); }
        if (d.pop < 2 || d.pop % nis != 0) { snprintf(buf, sizeof buf, "task %d: pop %d not a multiple of n_islands %d", t, d.pop, nis); return fail(EMTF_E_INVALID, buf); }
        if (d.pop > 1024) { snprintf(buf, sizeof buf, "task %d: pop %d > 1024", t, d.pop); return fail(EMTF_E_UNSUPPORTED, buf); }
        if (!(d.lo < d.hi) || !std::isfinite(d.lo) || !std::isfinite(d.hi)) { snprintf(buf, sizeof buf, "task %d: bad box", t); return fail(EMTF_E_INVALID, buf); }
        if (d.xform < EMTF_SHIFT || d.xform > EMTF_DIAG) { snprintf(buf, sizeof buf, "task %d: unknown xform", t); return fail(EMTF_E_INVALID, buf); }
        if (d.xform == EMTF_ROTATE && !d.rot) { snprintf(buf, sizeof buf, "task %d: rotation matrix missing", t); return fail(EMTF_E_INVALID, buf); }
        if (d.xform == EMTF_DIAG && !d.diag) { snprintf(buf, sizeof buf, "task %d: diagonal missing", t); return fail(EMTF_E_INVALID, buf); }
        // a NaN or infinity in the transform or the shift would make every fitness NaN
        auto finite = [](const float* v, size_t n) {
            for (size_t i = 0; v && i < n; ++i)
                if (!std::isfinite(v[i])) return false;
            return true;
        };
        if ((d.xform == EMTF_ROTATE && !finite(d.rot, (size_t)d.dim * d.dim)) ||
            (d.xform == EMTF_DIAG && !finite(d.diag, (size_t)d.dim)) || !finite(d.shift, (size_t)d.dim)) {
            snprintf(buf, sizeof buf, "task %d: non-finite value in rot / diag / shift", t);
            return fail(EMTF_E_INVALID, buf);
        }
        const int m = d.pop / nis;
        if (prm->algo == EMTF_GA_SBX_PM && (m % 2)) { snprintf(buf, sizeof buf, "task %d: GA needs an even island size", t); return fail(EMTF_E_INVALID, buf); }
        if (prm->transfer_interval > 0 && 2 * prm->transfer_count > d.pop) { snprintf(buf, sizeof buf, "task %d: transfer_count > pop/2", t); return fail(EMTF_E_INVALID, buf); }
        if (nis > 1 && prm->migrate_interval > 0 && 2 * prm->migrate_count > m) { snprintf(buf, sizeof buf, "task %d: migrate_count > island/2", t); return fail(EMTF_E_INVALID, buf); }
        if (d.dim > max_dim) max_dim = d.dim;
    }

    emtf_handle* h = new emtf_handle();
    h->T = n_tasks;
    h->max_dim = max_dim;
    emtf::EngineParams& p = h->p;
    p.algo = prm->algo;
    p.F = prm->F;
    p.CR = prm->CR;
    {
        const double c = std::floor((double)prm->CR * 65536.0);
        p.cr_threshold = (uint32_t)(c < 0.0 ? 0.0 : (c > 65536.0 ? 65536.0 : c));
    }
    p.eta_c = prm->eta_c;
    p.eta_m = prm->eta_m;
    p.pm = prm->pm;
    p.n_islands = nis;
    p.exchange_interval = prm->exchange_interval;
    p.migrate_interval = prm->migrate_interval;
    p.migrate_count = prm->migrate_count;
    p.transfer_interval = prm->transfer_interval;
    p.transfer_count = prm->transfer_interval > 0 ? prm->transfer_count : 0;
    p.transfer_cr = prm->transfer_cr;
    p.seed_lo = (uint32_t)prm->seed;
    p.seed_hi = (uint32_t)(prm->seed >> 32);
    p.polish_interval = prm->polish_interval;
    p.polish_steps = prm->polish_steps;
    p.polish_topk = prm->polish_topk;

    // bucket of every task: LDS kernel of the smallest KB whose population fits, else the wide kernel
    const bool ga = prm->algo == EMTF_GA_SBX_PM;
    const bool want_polish = prm->polish_interval > 0 && prm->polish_steps > 0 && prm->polish_topk > 0;
    std::vector<int> bucket_of(n_tasks), wide_nt(n_tasks, -1);
    std::vector<double> cost(n_tasks);
    for (int t = 0; t < n_tasks; ++t) {
        const emtf_task_desc& d = tasks[t];
        const int kb = pick_kb(d.dim);
        bool wide = kb < 0 || 4 * emtf_lds_floats_kb(d.pop, kb) > kMaxLdsBytes;
        const int geo = wide ? pick_nt(d.dim) : kb;
        if (wide && want_polish) {
            free_handle(h);
            char buf[160];
            snprintf(buf, sizeof buf, "task %d: gradient polish needs the population in LDS (dim %d, pop %d)", t, d.dim, d.pop);
            return fail(EMTF_E_UNSUPPORTED, buf);
        }
        int bi = -1;
        for (size_t k = 0; k < h->buckets.size(); ++k)
            if (h->buckets[k].wide == wide && h->buckets[k].geo == geo) bi = (int)k;
        if (bi < 0) {
            Bucket b;
            b.wide = wide;
            b.geo = geo;
            if (wide) {
                int (*wl)(int) = nullptr;
                b.fn = pick_wide(geo, ga, &wl);
                b.lds_bytes = 0;
            } else {
                b.fn = pick_launcher(geo, ga);
            }
            if (!b.fn) {
                free_handle(h);
                return fail(EMTF_E_UNSUPPORTED, "no kernel built for this dimension");
            }
            h->buckets.push_back(b);
            bi = (int)h->buckets.size() - 1;
        }
        bucket_of[t] = bi;
        if (wide) wide_nt[t] = geo;
        Bucket& b = h->buckets[bi];
        if (wide) {
            int (*wl)(int) = nullptr;
            pick_wide(geo, ga, &wl);
            b.lds_bytes = std::max(b.lds_bytes, 4 * wl(d.pop));
        } else {
            b.lds_bytes = std::max(b.lds_bytes, 4 * emtf_lds_floats_kb(d.pop, geo));
        }
        // relative cost per generation: objective work per gene + the rotation GEMM
        const int dk = wide ? 32 * geo : 8 * geo;
        cost[t] = func_cost(d.func) * d.pop * (double)d.dim + 2.0 * d.pop * (double)dk * dk / 64.0;
        b.cost += cost[t];
        ++b.count;
    }

    // variants: [0] as chosen above; [1] for an LDS-kernel bucket, the wide kernel for
    // its largest dimension, for a wide bucket, its 512-thread build; [2] the wide kernel
    // with Q in LDS where it fits (timed against [0] by the autotuner; not with polish,
    // whose kernel needs the LDS rows)
    // the persistent schedule runs one LDS-kernel launch unit, without polish or admission
    const bool persist = prm->schedule == EMTF_SCHEDULE_PERSISTENT;
    if (persist) {
        const char* why = nullptr;
        if (want_polish) why = "persistent schedule: gradient polish needs the epoch schedule";
        else if (online) why = "persistent schedule: online admission needs the epoch schedule";
        else if (h->buckets.size() != 1 || h->buckets[0].wide)
            why = "persistent schedule: every task must fall in one LDS-kernel bucket (one row geometry, "
                  "the population in LDS)";
        if (why) {
            free_handle(h);
            return fail(EMTF_E_UNSUPPORTED, why);
        }
    }
    const bool tune = prm->tune != 1 && !want_polish && !online && !persist;
    std::vector<int> alt_nt(h->buckets.size(), -1);
    for (size_t k = 0; k < h->buckets.size(); ++k) {
        Bucket& b = h->buckets[k];
        b.v[0].wide = b.wide;
        b.v[0].geo = b.geo;
        b.v[0].fn = b.fn;
        b.v[0].lds_bytes = b.lds_bytes;
        b.v[0].name = b.wide ? "wide NT" : "LDS KB";
        if (!tune) continue;
        int maxd = 0, maxp = 0;
        for (int t = 0; t < n_tasks; ++t)
            if (bucket_of[t] == (int)k) {
                maxd = std::max(maxd, tasks[t].dim);
                maxp = std::max(maxp, tasks[t].pop);
            }
        // the wide kernel with Q in LDS for row geometry nt, when its LDS fits
        auto add_wideq = [&](int nt) {
            int (*ql)(int) = nullptr;
            launch_fn qf = pick_wideq(nt, ga, &ql);
            if (!qf || 4 * ql(maxp) > kMaxLdsBytes) return;
            b.v[2].wide = true;
            b.v[2].geo = nt;
            b.v[2].fn = qf;
            b.v[2].lds_bytes = 4 * ql(maxp);
            b.v[2].name = "wide Q-in-LDS NT";
        };
        if (b.wide) {
            // wide bucket: the 512-thread build when some population spans more than four tiles
            if (maxp > 128) {
                b.v[1] = b.v[0];
                b.v[1].fn = pick_wide8(b.geo, ga);
                b.v[1].name = "wide x512 NT";
            }
            add_wideq(b.geo);
            continue;
        }
        // an LDS-kernel bucket: the wide kernel with Q in LDS always competes; the wide
        // kernel with Q in L2 only where LDS allows a single workgroup of the LDS kernel
        const int nt = pick_nt(maxd);
        if (nt <= 0) continue;
        add_wideq(nt);
        int (*wl)(int) = nullptr;
        launch_fn wf = 2 * b.lds_bytes > kMaxLdsBytes ? pick_wide(nt, ga, &wl) : nullptr;
        if (wf) {
            b.v[1].wide = true;
            b.v[1].geo = nt;
            b.v[1].fn = wf;
            b.v[1].lds_bytes = 4 * wl(maxp);
            b.v[1].name = "wide NT";
        }
        if (b.v[1].fn || b.v[2].fn) alt_nt[k] = nt;
    }
    for (int t = 0; t < n_tasks; ++t)
        if (alt_nt[bucket_of[t]] > 0) wide_nt[t] = alt_nt[bucket_of[t]];

    int64_t nX = 0, nF = 0, nQ = 0, nV = 0, nM = 0, nS = 0, nW = 0;
    h->tasks.resize(n_tasks);
    for (int t = 0; t < n_tasks; ++t) {
        const emtf_task_desc& d = tasks[t];
        const Bucket& b = h->buckets[bucket_of[t]];
        emtf::TaskDev& td = h->tasks[t];
        std::memset(&td, 0, sizeof td);
        td.func = kernel_func(d.func, prm->formulas);
        td.dim = d.dim;
        td.pop = d.pop;
        td.xform = d.xform;
        td.lo = d.lo;
        td.hi = d.hi;
        td.x_off = nX;
        td.fit_off = nF;
        td.q_off = nQ;
        td.v_off = nV;
        td.mb_off = nM;
        td.scr_off = nS;
        td.qw_off = nW;
        td.gen_base = online ? INT32_MAX : 0;  // not started until admitted
        td.gen_end = INT32_MAX;
        const int popP = (d.pop + 31) & ~31;
        nX += (int64_t)d.pop * d.dim;
        nF += d.pop;
        nQ += (int64_t)d.dim * d.dim;
        nV += d.dim;
        nM += (int64_t)p.transfer_count * d.dim;
        // scratch for every variant the task may run with
        int64_t scr = b.wide ? 0 : (int64_t)popP * (8 * b.geo + 4);
        if (wide_nt[t] > 0) {
            const int RS = 32 * wide_nt[t] + 4;
            scr = std::max(scr, 2 * (int64_t)popP * RS);  // population rows, trial rows
            nW += (int64_t)32 * wide_nt[t] * RS;
        }
        nS += scr;
    }
    for (size_t k = 0; k < h->buckets.size(); ++k)
        if (h->buckets[k].lds_bytes > kMaxLdsBytes) {
            free_handle(h);
            return fail(EMTF_E_UNSUPPORTED, "population tile exceeds LDS (pop x dim too large)");
        }
    if (want_polish) {
        h->polish = true;
        for (int t = 0; t < n_tasks; ++t) {
            Bucket& b = h->buckets[bucket_of[t]];
            if (b.polish_waves > 0) continue;
            polish_lds_fn plds = nullptr;
            if (!pick_polish(b.geo, &plds)) {
                free_handle(h);
                return fail(EMTF_E_UNSUPPORTED, "no polish kernel for this dimension");
            }
            int max_pop = 0;
            for (int u = 0; u < n_tasks; ++u)
                if (bucket_of[u] == bucket_of[t]) max_pop = std::max(max_pop, tasks[u].pop);
            // the elites go through in groups of 32 / cpe slots (all four waves on each
            // group): cpe columns per elite, the largest power of two <= 32 with every
            // elite in one group when they fit (up to 32), halved slots while LDS is short
            const int k = std::min(prm->polish_topk, max_pop);
            int slots = 1;  // a power of two >= min(k, 32)
            while (slots < std::min(k, 32)) slots *= 2;
            int cpe = 32 / slots;
            while (cpe < 32 && 4 * plds(max_pop, 4, cpe) > kMaxLdsBytes) cpe *= 2;
            if (4 * plds(max_pop, 4, cpe) > kMaxLdsBytes) {
                free_handle(h);
                return fail(EMTF_E_UNSUPPORTED, "polish staging exceeds LDS");
            }
            b.polish_waves = 4;
            b.polish_cpe = cpe;
            b.polish_lds_bytes = 4 * plds(max_pop, 4, cpe);
        }
    }
    // Launch units: each bucket's tasks, most expensive first.  With several buckets, one
    // carrying more than a quarter of the batch's cost is dealt round-robin (in its cost
    // order) into units of about that size, so that the caller's stream and up to three
    // more can be balanced; units go to the least-loaded stream, most expensive first.
    double total = 0.0;
    for (const Bucket& b : h->buckets) total += b.cost;
    std::vector<int> order;
    order.reserve(n_tasks);
    std::vector<Bucket> units;
    for (size_t r = 0; r < h->buckets.size(); ++r) {
        std::vector<int> ts;
        for (int t = 0; t < n_tasks; ++t)
            if (bucket_of[t] == (int)r) ts.push_back(t);
        std::stable_sort(ts.begin(), ts.end(), [&](int x, int y) { return cost[x] > cost[y]; });
        int k = 1;
        if (h->buckets.size() > 1) k = (int)std::ceil(h->buckets[r].cost / (total / 4.0) - 1e-9);
        k = std::max(1, std::min(k, (int)ts.size()));
        for (int u = 0; u < k; ++u) {
            Bucket ub = h->buckets[r];
            ub.offset = (int)order.size();
            ub.count = 0;
            ub.cost = 0.0;
            for (size_t i = u; i < ts.size(); i += k) {
                order.push_back(ts[i]);
                ub.cost += cost[ts[i]];
                ++ub.count;
            }
            units.push_back(ub);
        }
    }
    std::stable_sort(units.begin(), units.end(), [](const Bucket& x, const Bucket& y) { return x.cost > y.cost; });
    double load[4] = {0.0, 0.0, 0.0, 0.0};
    const int n_streams = std::min<int>(4, (int)units.size());
    for (Bucket& u : units) {
        int s = 0;
        for (int q = 1; q < n_streams; ++q)
            if (load[q] < load[s]) s = q;
        u.stream = s;
        load[s] += u.cost;
    }
    h->buckets = units;

    const int n_side = std::min<int>(3, (int)h->buckets.size() - 1);
    for (int k = 0; k < n_side; ++k) {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            free_handle(h);
            return fail(EMTF_E_DEVICE, "stream / event creation failed");
        }
        h->side.push_back(s);
        h->ev_join.push_back(ev);
    }
    if (n_side > 0 && hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess) {
        free_handle(h);
        return fail(EMTF_E_DEVICE, "event creation failed");
    }
    h->mailbox_stride = nM > 0 ? nM : 1;

    if (persist) {
        h->persist = pick_persist(h->buckets[0].geo, ga);
        if (!h->persist) {
            free_handle(h);
            return fail(EMTF_E_UNSUPPORTED, "persistent schedule: no kernel built for this row geometry");
        }
        const hipError_t ep = hipMalloc((void**)&h->d_persist, sizeof(int32_t) * (2 + (size_t)n_tasks + kPersistEpochs));
        if (ep != hipSuccess) {
            free_handle(h);
            return fail(EMTF_E_NOMEM, std::string("persistent schedule counters: ") + hipGetErrorString(ep));
        }
    }

    {
        // the device this handle lives on, and (once per device) the address of its fixed
        // staging array for emtf_get_elites_noalloc
        float* sp = nullptr;
        size_t cap = 0;
        hipError_t e0 = hipGetDevice(&h->device);
        if (e0 == hipSuccess) e0 = (hipError_t)emtf_scratch_resolve(h->device, &sp, &cap);
        if (e0 != hipSuccess) {
            free_handle(h);
            return fail(EMTF_E_DEVICE, std::string("device / staging array: ") + hipGetErrorString(e0));
        }
    }

    std::vector<float> Q((size_t)nQ, 0.0f), V((size_t)nV, 0.0f), TB((size_t)nV, 0.0f);
    for (int t = 0; t < n_tasks; ++t) {
        const emtf_task_desc& d = tasks[t];
        const emtf::TaskDev& td = h->tasks[t];
        float* q = Q.data() + td.q_off;
        for (int j = 0; j < d.dim; ++j) {
            if (d.xform == EMTF_ROTATE) std::memcpy(q + (size_t)j * d.dim, d.rot + (size_t)j * d.dim, sizeof(float) * d.dim);
            else q[(size_t)j * d.dim + j] = (d.xform == EMTF_DIAG) ? d.diag[j] : 1.0f;
        }
        if (d.shift) std::memcpy(V.data() + td.v_off, d.shift, sizeof(float) * d.dim);
        func_table(td.func, d.dim, TB.data() + td.v_off);
    }
    // wide tasks: Q rows in the kernels' [even | odd] row layout, zero padded to 32 NT rows
    std::vector<float> QW((size_t)nW, 0.0f);
    for (int t = 0; t < n_tasks; ++t) {
        const emtf::TaskDev& td = h->tasks[t];
        const int nt = wide_nt[t];
        if (nt < 0) continue;
        const int RS = 32 * nt + 4, HS = 16 * nt;
        const float* q = Q.data() + td.q_off;
        float* w = QW.data() + td.qw_off;
        for (int j = 0; j < td.dim; ++j)
            for (int d2 = 0; d2 < td.dim; ++d2) w[(size_t)j * RS + (d2 & 1) * HS + (d2 >> 1)] = q[(size_t)j * td.dim + d2];
    }

    auto alloc = [&](float** ptr, int64_t n) -> int {
        const hipError_t e = hipMalloc((void**)ptr, (size_t)(n > 0 ? n : 1) * sizeof(float));
        if (e != hipSuccess) return fail(EMTF_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        return EMTF_OK;
    };
    int rc = EMTF_OK;
    if ((rc = alloc(&h->d_X, nX)) || (rc = alloc(&h->d_fit, nF)) || (rc = alloc(&h->d_Q, nQ)) ||
        (rc = alloc(&h->d_shift, nV)) || (rc = alloc(&h->d_tab, nV)) || (rc = alloc(&h->d_mailbox, 2 * h->mailbox_stride)) ||
        (rc = alloc(&h->d_scratch, nS)) || (rc = alloc(&h->d_qw, nW))) {
        free_handle(h);
        return rc;
    }
    hipError_t e = hipMalloc((void**)&h->d_tasks, sizeof(emtf::TaskDev) * n_tasks);
    if (e != hipSuccess) { free_handle(h); return fail(EMTF_E_NOMEM, "hipMalloc(tasks)"); }
    e = hipMalloc((void**)&h->d_order, sizeof(int) * n_tasks);
    if (e != hipSuccess) { free_handle(h); return fail(EMTF_E_NOMEM, "hipMalloc(order)"); }
    e = hipMalloc((void**)&h->d_gen, sizeof(int32_t));
    if (e != hipSuccess) { free_handle(h); return fail(EMTF_E_NOMEM, "hipMalloc(generation counter)"); }
    h->online = online;
    if (online) {
        e = hipMalloc((void**)&h->d_donors, sizeof(int32_t) * ((size_t)n_tasks + 1));
        if (e != hipSuccess) { free_handle(h); return fail(EMTF_E_NOMEM, "hipMalloc(donor list)"); }
        // tasks not admitted yet read back as zero rows with NaN fitness
        e = hipMemset(h->d_X, 0, sizeof(float) * (size_t)std::max<int64_t>(nX, 1));
        if (e == hipSuccess) e = hipMemset(h->d_fit, 0xFF, sizeof(float) * (size_t)std::max<int64_t>(nF, 1));
        if (e != hipSuccess) { free_handle(h); return fail(EMTF_E_DEVICE, "initial state of an online handle"); }
    }
    // replayed epochs unless the caller asks for direct launches; the polish kernel's epochs
    // and those of online admission (whose launches change with the running tasks) are
    // launched directly
    h->use_graph = prm->schedule == EMTF_SCHEDULE_AUTO && !want_polish && !online;
#define CP(dst, src, bytes)                                                                    \
    do {                                                                                       \
        hipError_t e2 = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);                     \
        if (e2 != hipSuccess) { free_handle(h); return fail(EMTF_E_DEVICE, hipGetErrorString(e2)); } \
    } while (0)
    CP(h->d_tasks, h->tasks.data(), sizeof(emtf::TaskDev) * n_tasks);
    CP(h->d_order, order.data(), sizeof(int) * n_tasks);
    CP(h->d_Q, Q.data(), sizeof(float) * Q.size());
    CP(h->d_shift, V.data(), sizeof(float) * V.size());
    CP(h->d_tab, TB.data(), sizeof(float) * TB.size());
    if (nW > 0) CP(h->d_qw, QW.data(), sizeof(float) * QW.size());
#undef CP
    if (!online && (rc = launch(h, 0, 0, 1, nullptr)) != EMTF_OK) { free_handle(h); return rc; }
    e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) { free_handle(h); return fail(EMTF_E_DEVICE, std::string("init: ") + hipGetErrorString(e)); }
    if (tune && prm->tune == 0 && (rc = calibrate(h, nX, nF, order)) != EMTF_OK) { free_handle(h); return rc; }
    if (tune && prm->tune >= 2)
        for (Bucket& b : h->buckets) {
            const int v = prm->tune - 1;
            b.use = b.v[v].fn ? v : 0;
        }
    *out = h;
    return EMTF_OK;
}

// The persistent schedule: one launch per stretch of the run between trace points (at
// most kPersistEpochs epochs each), then the trace kernel where one falls.  Synchronous: the
// launches' error flag is read back at the end (a device-side wait that ran out).
static int run_persistent(emtf_handle* h, int generations, hipStream_t s) {
    const int TI = h->p.transfer_interval;
    const bool xfer = TI > 0 && h->T > 1 && h->p.transfer_count > 0;
    const int g_end = h->gen + generations;
    int32_t* err = h->d_persist;  // then the item counter, T stamps, the read counts
    int32_t* next = err + 1;
    const Bucket& b = h->buckets[0];
    emtf::EpochArgs a{};
    a.tasks = h->d_tasks;
    a.n_tasks = h->T;
    a.X = h->d_X;
    a.fit = h->d_fit;
    a.Q = h->d_Q;
    a.shift = h->d_shift;
    a.tab = h->d_tab;
    a.mailbox = h->d_mailbox;
    a.mailbox_stride = h->mailbox_stride;
    a.scratch = h->d_scratch;
    a.Qw = h->d_qw;
    a.p = h->p;
    int e = emtf_launch_fill_i32(err, 1, 0, s);
    if (e != 0) return fail(EMTF_E_DEVICE, std::string("persistent counters: ") + hipGetErrorString((hipError_t)e));
    while (h->gen < g_end) {
        int g1 = g_end;
        const int TR = h->trace_points < h->trace_max ? h->trace_interval : 0;
        if (TR > 0) {
            const int nb = h->trace_g0 + ((h->gen - h->trace_g0) / TR + 1) * TR;
            if (nb < g1) g1 = nb;
        }
        const bool trace_here = TR > 0 && (g1 - h->trace_g0) % TR == 0;
        int n_ep = 1;
        if (xfer) {
            // at most kPersistEpochs epochs, and items (epochs x tasks) within 2^30
            const int max_ep = std::max(1, std::min(kPersistEpochs, (1 << 30) / h->T));
            const int64_t lim = ((int64_t)(h->gen / TI) + max_ep) * TI;
            if (lim < g1) g1 = (int)lim;
            n_ep = (g1 - 1) / TI - h->gen / TI + 1;  // cut at the boundaries inside (gen, g1)
        }
        e = emtf_launch_fill_i32(next, 1 + h->T + n_ep, 0, s);
        if (e != 0) return fail(EMTF_E_DEVICE, std::string("persistent counters: ") + hipGetErrorString((hipError_t)e));
        emtf::PersistArgs pa{};
        pa.G0 = h->gen;
        pa.G1 = g1;
        pa.n_epochs = n_ep;
        pa.next = next;
        pa.err = err;
        pa.stamp = next + 1;
        pa.reads = next + 1 + h->T;
        pa.wait_ticks = kPersistWaitTicks;
        const int blocks = n_ep * h->T;  // one workgroup per item
        const int rc = h->persist(&a, &pa, h->d_order + b.offset, blocks, b.v[0].lds_bytes, s);
        if (rc != 0) return fail(EMTF_E_DEVICE, std::string("persistent kernel launch: ") + hipGetErrorString((hipError_t)rc));
        h->gen = g1;
        h->gen_valid = false;
        if (trace_here) {
            e = emtf_launch_trace(h->d_tasks, h->T, h->d_fit, h->d_trace + (int64_t)h->trace_points * h->T, nullptr, s);
            if (e != 0) return fail(EMTF_E_DEVICE, std::string("trace: ") + hipGetErrorString((hipError_t)e));
            ++h->trace_points;
        }
    }
    int32_t flag = 0;
    HIP_OR_FAIL(hipMemcpyAsync(&flag, err, sizeof flag, hipMemcpyDeviceToHost, s));
    HIP_OR_FAIL(hipStreamSynchronize(s));
    if (flag) return fail(EMTF_E_DEVICE, "persistent schedule: a device-side wait ran out");
    return EMTF_OK;
}

int emtf_run(emtf_handle* h, int32_t generations, void* stream) {
    g_err.clear();
    if (!h) return fail(EMTF_E_INVALID, "handle is NULL");
    if (generations < 0) return fail(EMTF_E_INVALID, "generations < 0");
    hipStream_t s = (hipStream_t)stream;
    if (const int rc = use_stream(h, s); rc != EMTF_OK) return rc;
    if (h->persist) return run_persistent(h, generations, s);
    const int TI = h->p.transfer_interval;
    const bool xfer = TI > 0 && h->T > 1 && h->p.transfer_count > 0;
    const int g_end = h->gen + generations;
    const int PI = h->polish ? h->p.polish_interval : 0;
    while (h->gen < g_end) {
        int g1 = g_end;
        if (xfer) {
            const int nb = (h->gen / TI + 1) * TI;
            if (nb < g1) g1 = nb;
        }
        if (PI > 0) {
            const int nb = (h->gen / PI + 1) * PI;
            if (nb < g1) g1 = nb;
        }
        // online: an epoch also ends where an admitted task finishes
        const auto ne = std::upper_bound(h->ends.begin(), h->ends.end(), h->gen);
        if (ne != h->ends.end() && *ne < g1) g1 = *ne;
        const int TR = h->trace_points < h->trace_max ? h->trace_interval : 0;
        if (TR > 0) {
            const int nb = h->trace_g0 + ((h->gen - h->trace_g0) / TR + 1) * TR;
            if (nb < g1) g1 = nb;
        }
        const bool trace_here = TR > 0 && (g1 - h->trace_g0) % TR == 0;
        // a replayable epoch: no trace point or polish at its end and, with transfer, one
        // whole transfer interval (so every replay ends on a boundary and publishes)
        const bool replay = h->use_graph && !trace_here && !polish_at(h, g1) &&
                            (!xfer || (h->gen % TI == 0 && g1 - h->gen == TI));
        int rc;
        if (replay) {
            rc = replay_epoch(h, h->gen, g1 - h->gen, s);
        } else {
            rc = launch(h, h->gen, g1, 0, s);
            h->gen_valid = false;
        }
        if (rc != EMTF_OK) return rc;
        h->gen = g1;
        if (trace_here) {
            const int e = emtf_launch_trace(h->d_tasks, h->T, h->d_fit, h->d_trace + (int64_t)h->trace_points * h->T,
                                            nullptr, s);
            if (e != 0) return fail(EMTF_E_DEVICE, std::string("trace: ") + hipGetErrorString((hipError_t)e));
            ++h->trace_points;
        }
    }
    return EMTF_OK;
}

int emtf_admit(emtf_handle* h, int32_t first, int32_t count, int32_t generations, void* stream) {
    g_err.clear();
    if (!h) return fail(EMTF_E_INVALID, "handle is NULL");
    if (!h->online) return fail(EMTF_E_INVALID, "the handle was created with admission AT_CREATE");
    if (first < 0 || count < 1 || first > h->T - count) return fail(EMTF_E_INVALID, "task range out of bounds");
    if (generations < 0 || generations > INT32_MAX - 1 - h->gen) return fail(EMTF_E_INVALID, "bad generation count");
    const int TI = h->p.transfer_interval;
    if (TI > 0 && h->gen % TI != 0) return fail(EMTF_E_INVALID, "admission off a transfer boundary");
    for (int t = first; t < first + count; ++t)
        if (h->tasks[t].gen_base != INT32_MAX) return fail(EMTF_E_INVALID, "task admitted before");
    hipStream_t s = (hipStream_t)stream;
    if (const int rc = use_stream(h, s); rc != EMTF_OK) return rc;
    for (int t = first; t < first + count; ++t) {
        h->tasks[t].gen_base = h->gen;
        h->tasks[t].gen_end = h->gen + generations;
    }
    h->ends.insert(std::upper_bound(h->ends.begin(), h->ends.end(), h->gen + generations), h->gen + generations);
    // the descriptors' host copies stay as written for the life of the handle
    HIP_OR_FAIL(hipMemcpyAsync(h->d_tasks + first, h->tasks.data() + first, sizeof(emtf::TaskDev) * count,
                               hipMemcpyHostToDevice, s));
    return launch(h, h->gen, h->gen, 1, s);
}

int emtf_trace_start(emtf_handle* h, int32_t interval, int32_t max_points, void* stream) {
    g_err.clear();
    if (!h) return fail(EMTF_E_INVALID, "handle is NULL");
    if (interval < 0) return fail(EMTF_E_INVALID, "interval < 0");
    if (interval == 0) {
        h->trace_interval = 0;
        return EMTF_OK;
    }
    if (max_points < 1) return fail(EMTF_E_INVALID, "max_points < 1");
    // the previous run's stream first: its trace kernels may still write the old buffer
    HIP_OR_FAIL(hipStreamSynchronize(h->stream));
    if (const int rc = use_stream(h, (hipStream_t)stream); rc != EMTF_OK) return rc;
    (void)hipFree(h->d_trace);
    (void)hipFree(h->d_trace_worst);
    h->d_trace = h->d_trace_worst = nullptr;
    h->trace_interval = h->trace_max = h->trace_points = 0;
    HIP_OR_FAIL(hipMalloc((void**)&h->d_trace, sizeof(float) * (size_t)max_points * h->T));
    HIP_OR_FAIL(hipMalloc((void**)&h->d_trace_worst, sizeof(float) * (size_t)h->T));
    const int e = emtf_launch_trace(h->d_tasks, h->T, h->d_fit, h->d_trace, h->d_trace_worst, h->stream);
    if (e != 0) return fail(EMTF_E_DEVICE, std::string("trace: ") + hipGetErrorString((hipError_t)e));
    h->trace_interval = interval;
    h->trace_max = max_points;
    h->trace_points = 1;
    h->trace_g0 = h->gen;
    return EMTF_OK;
}

int emtf_trace(emtf_handle* h, int32_t* n_points, float* best, float* initial_worst) {
    g_err.clear();
    if (!h || !n_points) return fail(EMTF_E_INVALID, "NULL argument");
    HIP_OR_FAIL(hipStreamSynchronize(h->stream));
    *n_points = h->trace_points;
    if (h->trace_points == 0) return EMTF_OK;
    if (best)
        HIP_OR_FAIL(hipMemcpy(best, h->d_trace, sizeof(float) * (size_t)h->trace_points * h->T, hipMemcpyDeviceToHost));
    if (initial_worst)
        HIP_OR_FAIL(hipMemcpy(initial_worst, h->d_trace_worst, sizeof(float) * (size_t)h->T, hipMemcpyDeviceToHost));
    return EMTF_OK;
}

int emtf_sync(emtf_handle* h) {
    if (!h) return fail(EMTF_E_INVALID, "handle is NULL");
    HIP_OR_FAIL(hipStreamSynchronize(h->stream));
    return EMTF_OK;
}

int emtf_generation(const emtf_handle* h, int32_t* generation) {
    if (!h || !generation) return fail(EMTF_E_INVALID, "NULL argument");
    *generation = h->gen;
    return EMTF_OK;
}

int emtf_get_population(emtf_handle* h, int32_t task, float* x, float* fitness) {
    if (!h || task < 0 || task >= h->T) return fail(EMTF_E_INVALID, "bad handle or task index");
    HIP_OR_FAIL(hipStreamSynchronize(h->stream));
    const emtf::TaskDev& td = h->tasks[task];
    if (x) HIP_OR_FAIL(hipMemcpy(x, h->d_X + td.x_off, sizeof(float) * (size_t)td.pop * td.dim, hipMemcpyDeviceToHost));
    if (fitness) HIP_OR_FAIL(hipMemcpy(fitness, h->d_fit + td.fit_off, sizeof(float) * (size_t)td.pop, hipMemcpyDeviceToHost));
    return EMTF_OK;
}

int emtf_set_population(emtf_handle* h, int32_t task, const float* x, const float* fitness) {
    if (!h || task < 0 || task >= h->T) return fail(EMTF_E_INVALID, "bad handle or task index");
    HIP_OR_FAIL(hipStreamSynchronize(h->stream));
    const emtf::TaskDev& td = h->tasks[task];
    if (x) HIP_OR_FAIL(hipMemcpy(h->d_X + td.x_off, x, sizeof(float) * (size_t)td.pop * td.dim, hipMemcpyHostToDevice));
    if (fitness) HIP_OR_FAIL(hipMemcpy(h->d_fit + td.fit_off, fitness, sizeof(float) * (size_t)td.pop, hipMemcpyHostToDevice));
    return EMTF_OK;
}

int emtf_get_best(emtf_handle* h, float* fitness, float* x_best, int32_t x_stride) {
    if (!h) return fail(EMTF_E_INVALID, "handle is NULL");
    if (x_best && x_stride < h->max_dim) return fail(EMTF_E_INVALID, "x_stride < max dim");
    // the first of the device ranking by (fitness, index) with NaN last: the lowest index
    // among the smallest fitness values
    return emtf_get_elites(h, 1, fitness, x_best, x_stride);
}

static int objective_call(int32_t func, int32_t dim, int32_t n, const float* x, const float* rot, const float* shift,
                          const float* gf, float* f, float* gx, void* stream, int adjoint) {
    g_err.clear();
    if (func < EMTF_SPHERE || func > EMTF_ELLIPTIC) return fail(EMTF_E_INVALID, "unknown func");
    if (dim < 1) return fail(EMTF_E_INVALID, "dim < 1");
    if (dim > 256) return fail(EMTF_E_UNSUPPORTED, "dim > 256 not built");
    if (adjoint && dim > 128) return fail(EMTF_E_UNSUPPORTED, "adjoint: dim > 128 not built");
    if (n < 0 || !x || (adjoint ? (!gf || !gx) : !f)) return fail(EMTF_E_INVALID, "NULL buffer or n < 0");
    if (n == 0) return EMTF_OK;
    emtf::ObjArgs a{};
    a.func = func;
    a.dim = dim;
    a.n = n;
    a.has_rot = rot != nullptr;
    a.x = x;
    a.rot = rot;
    a.shift = shift;
    a.gf = gf;
    a.f = f;
    a.gx = gx;
    func_table(func, dim, a.tab);
    int rc = -1;
    if (dim > 128) {
        switch (pick_nt(dim)) {
        case 5: rc = emtf_launch_objective_nt5_fwd(&a, (hipStream_t)stream); break;
        case 6: rc = emtf_launch_objective_nt6_fwd(&a, (hipStream_t)stream); break;
        case 7: rc = emtf_launch_objective_nt7_fwd(&a, (hipStream_t)stream); break;
        case 8: rc = emtf_launch_objective_nt8_fwd(&a, (hipStream_t)stream); break;
        }
    } else {
        switch (pick_kb(dim)) {
        case 2: rc = emtf_launch_objective_kb2(&a, adjoint, (hipStream_t)stream); break;
        case 4: rc = emtf_launch_objective_kb4(&a, adjoint, (hipStream_t)stream); break;
        case 7: rc = emtf_launch_objective_kb7(&a, adjoint, (hipStream_t)stream); break;
        case 8: rc = emtf_launch_objective_kb8(&a, adjoint, (hipStream_t)stream); break;
        case 12: rc = emtf_launch_objective_kb12(&a, adjoint, (hipStream_t)stream); break;
        case 13: rc = emtf_launch_objective_kb13(&a, adjoint, (hipStream_t)stream); break;
        case 16: rc = emtf_launch_objective_kb16(&a, adjoint, (hipStream_t)stream); break;
        }
    }
    if (rc != 0) return fail(EMTF_E_DEVICE, std::string("objective kernel launch: ") + hipGetErrorString((hipError_t)rc));
    return EMTF_OK;
}

int emtf_objective_fwd(int32_t func, int32_t dim, int32_t n, const float* x, const float* rot, const float* shift,
                       float* f, void* stream) {
    return objective_call(func, dim, n, x, rot, shift, nullptr, f, nullptr, stream, 0);
}

int emtf_objective_adjoint(int32_t func, int32_t dim, int32_t n, const float* x, const float* rot, const float* shift,
                           const float* gf, float* gx, void* stream) {
    return objective_call(func, dim, n, x, rot, shift, gf, nullptr, gx, stream, 1);
}

int emtf_eval_draws(uint64_t seed, int32_t n, const uint32_t* addrs, uint32_t* out, void* stream) {
    g_err.clear();
    if (n < 0) return fail(EMTF_E_INVALID, "n < 0");
    if (n == 0) return EMTF_OK;
    if (!addrs || !out) return fail(EMTF_E_INVALID, "addrs or out is NULL");
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0) return fail(EMTF_E_INVALID, "out is not 16-byte aligned");
    const int rc = emtf_launch_draws(seed, n, addrs, out, (hipStream_t)stream);
    if (rc != 0) return fail(EMTF_E_DEVICE, std::string("draws: ") + hipGetErrorString((hipError_t)rc));
    return EMTF_OK;
}

int emtf_get_elites(emtf_handle* h, int32_t k, float* fitness, float* x, int32_t x_stride) {
    g_err.clear();
    if (!h) return fail(EMTF_E_INVALID, "handle is NULL");
    if (k < 1) return fail(EMTF_E_INVALID, "k < 1");
    if (x && x_stride < h->max_dim) return fail(EMTF_E_INVALID, "x_stride < max dim");
    for (const emtf::TaskDev& td : h->tasks)
        if (k > td.pop) return fail(EMTF_E_INVALID, "k exceeds a task's population");
    const size_t nx = x ? (size_t)h->T * k * x_stride : 0, nf = (size_t)h->T * k;
    hipError_t e = hipSuccess;
    if (nf + nx > h->elite_cap) {
        e = hipStreamSynchronize(h->stream);  // a previous export may still read the old buffer
        (void)hipFree(h->d_elite);
        h->d_elite = nullptr;
        h->elite_cap = 0;
        if (e == hipSuccess) e = hipMalloc((void**)&h->d_elite, (nf + nx) * sizeof(float));
        if (e == hipSuccess) h->elite_cap = nf + nx;
    }
    float* df = h->d_elite;
    float* dx = x ? h->d_elite + nf : nullptr;
    if (e == hipSuccess)
        e = (hipError_t)emtf_launch_elites(h->d_tasks, h->T, h->d_X, h->d_fit, k, x_stride, dx, df, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && fitness) e = hipMemcpy(fitness, df, nf * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && x) e = hipMemcpy(x, dx, nx * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(EMTF_E_DEVICE, std::string("elite export: ") + hipGetErrorString(e));
    return EMTF_OK;
}

int emtf_get_elites_noalloc(emtf_handle* h, int32_t k, float* fitness, float* x, int32_t x_stride) {
    g_err.clear();
    if (!h) return fail(EMTF_E_INVALID, "handle is NULL");
    if (k < 1) return fail(EMTF_E_INVALID, "k < 1");
    if (x && x_stride < h->max_dim) return fail(EMTF_E_INVALID, "x_stride < max dim");
    for (const emtf::TaskDev& td : h->tasks)
        if (k > td.pop) return fail(EMTF_E_INVALID, "k exceeds a task's population");
    float* base = nullptr;
    size_t cap = 0;
    // resolved at create for this device: a table lookup here, no allocation
    hipError_t e = (hipError_t)emtf_scratch_resolve(h->device, &base, &cap);
    if (e != hipSuccess) return fail(EMTF_E_DEVICE, std::string("staging array: ") + hipGetErrorString(e));
    const size_t nx = x ? (size_t)h->T * k * x_stride : 0, nf = (size_t)h->T * k;
    if (nf + nx > cap) {
        char buf[200];
        snprintf(buf, sizeof buf, "elite export needs %zu floats of staging; the fixed array holds %zu "
                 "(n_tasks * k * (x_stride + 1) at most)", nf + nx, cap);
        return fail(EMTF_E_INVALID, buf);
    }
    float* df = base;
    float* dx = x ? base + nf : nullptr;
    e = (hipError_t)emtf_launch_elites(h->d_tasks, h->T, h->d_X, h->d_fit, k, x_stride, dx, df, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && fitness) e = hipMemcpy(fitness, df, nf * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && x) e = hipMemcpy(x, dx, nx * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(EMTF_E_DEVICE, std::string("elite export: ") + hipGetErrorString(e));
    return EMTF_OK;
}

int emtf_destroy(emtf_handle* h) {
    if (!h) return EMTF_OK;
    (void)hipStreamSynchronize(h->stream);
    free_handle(h);
    return EMTF_OK;
}

}  // extern "C"
