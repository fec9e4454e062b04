// island.hip -- the island-model evolutionary-multitasking kernel for gfx950.
//
// One workgroup (4 wave64s) owns one task for one epoch of generations.  The
// task's population, its transform matrix and its per-dimension tables live in
// LDS for the whole epoch; HBM is touched only to load and store the state at
// the epoch boundaries and for the elite mailbox that carries inter-task
// transfer.  Per generation and per 32-individual tile a wave
//   1. draws its Philox numbers in registers (no RNG state in memory),
//   2. gathers parent/mate rows from LDS and forms the trial vectors,
//   3. computes z = Q (u - o) with v_mfma_f32_32x32x2_f32: B = the trial tile
//      (one individual per column, one gene parity per lane half), A = Q,
//      so each lane ends up holding 16 rows of z for its own individual,
//   4. evaluates the objective on those rows and combines the two lane halves
//      with one cross-half shuffle,
//   5. applies the greedy one-to-one selection in LDS.
//
// Reference operators restated (file:line of proposed model/):
//   GeneticOpera SG-EMTF.cu:1289-1426 / SG-EMTF_DE.cu:1241-1390 (the island loop),
//   CrossOver/Mutation/Selection SG-EMTF_DE.cu:474-705, SG-EMTF.cu:548-831,
//   evaluation SG-EMTF_DE.cu:834-1069, IslandMigration SG-EMTF_DE.cu:1093-1170,
//   SharedToGlobal/SelectFromBlocks SG-EMTF.cu:856-896,1220-1260,
//   knowledgeTransfer SG-EMTF.cu:1514-1591, pop_init SG-EMTF.cu:146-159.
//
// LDS layout of a row (individual i, or row j of Q): the even genes
// 0,2,4,.. in slots [0, HS) and the odd genes in [HS, 2 HS), HS = 4 KB, so a
// lane half streams its parity with 16-byte loads; RS = 2 HS + 4 is the row
// stride.  Slots past the task's dimension hold zeros, which makes every MFMA
// step and every gene block unconditional: padded steps add +0 products.
#include <hip/hip_runtime.h>
#include "emtf_math.h"
#include "emtf_types.h"

namespace emtf {

typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;

// Row geometry fixed at compile time: KB blocks of 4 genes per parity half.
template <int KB>
struct Shape {
    static constexpr int HS = 4 * KB;          // slots per parity half
    static constexpr int RS = 2 * HS + 4;      // LDS row stride (floats)
    static constexpr int NT = (KB + 3) / 4;    // 32-row output tiles (dims <= 32 NT)
};

struct Geo {
    int D, Dh, pop, popP, ntiles, m;
    FastMod modD, modM;  // jrand = w mod D, island-local index = w mod m
};

struct Lds {
    float* X;    // popP x RS
    float* Q;    // 32 NT x RS
    float* o;    // RS
    float* tab;  // 32 NT
    float* fit;  // popP
    float* ftmp; // popP
    int* ord;    // popP
    int* aux;    // popP
};

__host__ __device__ inline int lds_floats_kb(int pop, int kb) {
    const int RS = 8 * kb + 4, NT = (kb + 3) / 4, popP = (pop + 31) & ~31;
    return popP * RS + 32 * NT * RS + RS + 32 * NT + 4 * popP;
}

__device__ __forceinline__ float4 lds4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ float f4(const float4& v, int e) { return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w)); }

// v clamped to [lo, hi] (hi if v > hi, else lo if v < lo, else v: a NaN stays NaN), as
// two selects.
__device__ __forceinline__ float clampf(float v, float lo, float hi) {
    const float t = v < lo ? lo : v;
    return v > hi ? hi : t;
}

// The box clamp of a trial gene as one min and one max: equal to clampf for every
// non-NaN value (trial genes are always finite) except a signed-zero tie with a bound
// of zero, which the engine never produces (a -0 gene needs a -0 parent, and the
// initial population of a box starting at 0 holds none).
__device__ __forceinline__ float clamp_box(float v, float lo, float hi) {
    return __builtin_fminf(__builtin_fmaxf(v, lo), hi);
}

// A replayed epoch (hipGraph) starts at the device generation counter: its launch
// arguments were frozen when the graph was captured, the counter advances every replay.
__device__ __forceinline__ void rebase_epoch(EpochArgs& a) {
    if (a.gen_dev) {
        const int g0 = *a.gen_dev;
        a.g1 = g0 + (a.g1 - a.g0);
        a.g0 = g0;
    }
}

// Online admission: a task takes part in the epochs that lie inside its life and is
// initialised by the launch at its admission (every task of a handle created without
// online admission: gen_base 0, gen_end INT32_MAX).
__device__ __forceinline__ bool runs_epoch(const EpochArgs& a, const TaskDev& td) {
    return a.init ? td.gen_base == a.g0 : (td.gen_base <= a.g0 && a.g1 <= td.gen_end);
}

// Whether the task receives elites at the start of this epoch: a transfer boundary that
// is not its own start, with another donor to draw.
// (g0: the epoch's first generation, a.g0 but for the persistent schedule's items)
__device__ __forceinline__ bool transfer_now(const EpochArgs& a, const TaskDev& td, int g0) {
    const int TI = a.p.transfer_interval;
    if (a.init || TI <= 0 || a.p.transfer_count <= 0 || g0 % TI != 0 || g0 <= td.gen_base) return false;
    return a.donors ? *a.n_donors > 1 : a.n_tasks > 1;
}
__device__ __forceinline__ bool transfer_now(const EpochArgs& a, const TaskDev& td) { return transfer_now(a, td, a.g0); }

// The donor of task `task` at a transfer event: the k-th of the other donors in ascending
// id order, k = w mod (n - 1) -- every task of the handle, or the online donor list.
__device__ __forceinline__ int pick_donor(const EpochArgs& a, int task, uint32_t w) {
    if (!a.donors) {
        const int dn = (int)(w % (uint32_t)(a.n_tasks - 1));
        return dn >= task ? dn + 1 : dn;
    }
    const int n = *a.n_donors;
    const int k = (int)(w % (uint32_t)(n - 1));
    int lo = 0, hi = n - 1;  // the task's own position in the list (it is a donor itself)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a.donors[mid] < task) lo = mid + 1;
        else hi = mid;
    }
    return a.donors[k + (k >= lo ? 1 : 0)];
}

// ---------------------------------------------------------------------------
// Objective of the lane's individual from its MFMA accumulators.
// Lane (col, h) holds z_j for j = 32 jt + 8 (r >> 2) + 4 h + (r & 3); the two
// halves accumulate the even / odd blocks of four rows in increasing j and
// are combined once.
// ZPAD: the caller's Q rows past D are zero, so z_j = +0 there; the objectives whose
// term at z = +0 is +0 (whose sums it leaves as they are, bit for bit: they start at
// +0 and never reach -0) and costs what any other row's does then add those rows
// unmasked.  (Weierstrass keeps the mask: at z = 0 its term takes the small-|z| branch,
// which would diverge from the other rows of the same slot.)
// ---------------------------------------------------------------------------
template <int FUNC>
constexpr bool zero_term_at_zero() {
    return FUNC == F_SPHERE || FUNC == F_ELLIPTIC || FUNC == F_ACKLEY || FUNC == F_RASTRIGIN ||
           FUNC == F_GRIEWANK || FUNC == F_SCHWEFEL || FUNC == F_REF_RASTRIGIN;
}

// The objective's terms and final combination in split form, for the polish's line search
// (terms of a row tile computed apart from the running sums): the same operations, in the
// same order, as objective() below, which keeps them inline for the epoch kernels' code.
//
// Rosenbrock's term of row j (z = z_j, zn = z_{j+1}); it counts when row j + 1 exists.
__device__ __forceinline__ float rosen_val(float z, float zn) {
    const float zz = z * z;
    const float a = zz - zn;
    const float b = z - 1.0f;
    float t1 = 100.0f * a;
    t1 = t1 * a;
    const float b2 = b * b;
    return t1 + b2;
}

// The term values of row j of every textbook objective but Rosenbrock (v2: Ackley's
// cosine term, Griewank's 2 sin^2); the Weierstrass term only for rows j < Dm (0 otherwise).
template <int FUNC>
__device__ __forceinline__ void row_vals(float z, int j, int Dm, const float* tab, float& v1, float& v2) {
    static_assert(FUNC <= F_ELLIPTIC, "textbook objectives only");
    if constexpr (FUNC == F_SPHERE) {
        v1 = z * z;
    } else if constexpr (FUNC == F_ELLIPTIC) {
        const float zz = z * z;
        v1 = tab[j] * zz;
    } else if constexpr (FUNC == F_ACKLEY) {
        v1 = z * z;
        v2 = sinpi_sq(z);
    } else if constexpr (FUNC == F_RASTRIGIN) {
        const float zz = z * z;
        const float s20 = 20.0f * sinpi_sq(z);
        v1 = zz + s20;
    } else if constexpr (FUNC == F_GRIEWANK) {
        v1 = z * z;
        v2 = 2.0f * sinpi_sq(z * tab[j]);
    } else if constexpr (FUNC == F_WEIERSTRASS) {
        v1 = 0.0f;
        if (j < Dm) v1 = weier_term(z);
    } else if constexpr (FUNC == F_SCHWEFEL) {
        const float sq = sqrtf(__builtin_fabsf(z));
        const float tv = sq * 0.159154937f;
        v1 = z * sin2pi(tv);
    }
}

// Adds one row's term values into the lane's sums (s2: Ackley's cosine sum, Griewank's
// running 1 - prod cos) when the row counts.
template <int FUNC>
__device__ __forceinline__ void row_sum(float v1, float v2, bool in, float& s1, float& s2) {
    if (!in) return;
    s1 = s1 + v1;
    if constexpr (FUNC == F_ACKLEY) {
        s2 = s2 + v2;
    } else if constexpr (FUNC == F_GRIEWANK) {
        const float om = 1.0f - s2;
        const float aq = v2 * om;
        s2 = s2 + aq;
    }
}

// The fitness from the two halves' sums (lane h holds its half's s1, s2; the other
// half's come over the cross-half shuffle), textbook objectives.
template <int FUNC>
__device__ __forceinline__ float objective_tail(float s1, float s2, int D, const float* tab, int h) {
    static_assert(FUNC <= F_ELLIPTIC, "textbook objectives only");
    const float o1 = __shfl_xor(s1, 32);
    const float S = s1 + o1;  // A + B (commutative)
    const float Df = (float)D;
    if constexpr (FUNC == F_ACKLEY) {
        const float o2 = __shfl_xor(s2, 32);
        const float C = s2 + o2;
        const float q = S / Df;
        const float r = sqrtf(q);
        const float a = -0.2f * r;
        const float t1 = -20.0f * expm1_f(a);
        const float c2 = C / Df;
        const float b = -2.0f * c2;
        const float t2 = 2.71828175f * expm1_f(b);
        return t1 - t2;
    } else if constexpr (FUNC == F_GRIEWANK) {
        const float o2 = __shfl_xor(s2, 32);
        const float qA = h ? o2 : s2;
        const float qB = h ? s2 : o2;
        const float om = 1.0f - qA;
        const float tq = qB * om;
        const float Qv = qA + tq;
        const float s4 = S / 4000.0f;
        return s4 + Qv;
    } else if constexpr (FUNC == F_SCHWEFEL) {
        // the constant from the task table: 418.982887 (textbook) or 418.9829 (reference)
        const float cd = tab[0] * Df;
        return cd - S;
    } else {
        return S;
    }
}

template <int NT, int FUNC, bool SKIP = false, bool ZPAD = false>
__device__ __forceinline__ float objective(const f16v (&acc)[NT], int D, const float* tab, int h) {
    // the row mask's bound: with ZPAD one no row reaches
    const int Dm = (ZPAD && zero_term_at_zero<FUNC>()) ? 0x7FFFFFFF : D;
    float s1 = 0.0f, s2 = (FUNC == F_REF_GRIEWANK) ? 1.0f : 0.0f;
    if constexpr (FUNC == F_ROSENBROCK) {
        float nb[NT][4];
#pragma unroll
        for (int jt = 0; jt < NT; ++jt)
#pragma unroll
            for (int g = 0; g < 4; ++g) nb[jt][g] = __shfl_xor(acc[jt][4 * g], 32);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int j = 32 * jt + 8 * (r >> 2) + 4 * h + (r & 3);
                const int g = r >> 2;
                float zn;
                if ((r & 3) != 3) zn = acc[jt][r + 1];
                else if (g < 3) zn = h ? nb[jt][g + 1] : nb[jt][g];
                else zn = h ? ((jt + 1 < NT) ? nb[(jt + 1) % NT][0] : 0.0f) : nb[jt][g];
                const float z = acc[jt][r];
                const float zz = z * z;
                const float a = zz - zn;
                const float b = z - 1.0f;
                float t1 = 100.0f * a;
                t1 = t1 * a;
                const float b2 = b * b;
                const float v = t1 + b2;
                if (j + 1 < D) s1 = s1 + v;
            }
    } else if constexpr (FUNC == F_WEIERSTRASS && SKIP && !ZPAD) {
        // The wide kernels (SKIP, rows not zero-padded): four rows at a time, their terms
        // by the branch-free path (independent chains the scheduler interleaves), the rare
        // arguments redone by the full term, then the row sums in row order as below (s1 is
        // never -0, so adding +0 for a masked row leaves it).  The LDS epoch kernel keeps
        // one row at a time: there the interleaved chains take issue slots from the tasks
        // that share the CU and the mixed epoch is slower.
#pragma unroll
        for (int jt = 0; jt < NT; ++jt)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                if (32 * jt + 8 * g >= D) continue;
                float v[4];
                bool rare = false;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float a = __builtin_fabsf(acc[jt][4 * g + e]);
                    const bool lo = a >= 2.32830644e-10f, hi = a >= 8388608.0f;
                    v[e] = weier_core(a, lo && !hi);
                    rare = rare || (!lo && !hi && a != 0.0f);  // NaN or 0 < |z| < 2^-32
                }
                if (rare) {
                    // one inlined copy of the full term: a rolled loop, the row picked by selects
#pragma unroll 1
                    for (int e = 0; e < 4; ++e) {
                        float z = acc[jt][4 * g];
                        z = e == 1 ? acc[jt][4 * g + 1] : z;
                        z = e == 2 ? acc[jt][4 * g + 2] : z;
                        z = e == 3 ? acc[jt][4 * g + 3] : z;
                        const float a = __builtin_fabsf(z);
                        if (z != z || (a > 0.0f && a < 2.32830644e-10f)) {
                            const float r = weier_term(z);
#pragma unroll
                            for (int f = 0; f < 4; ++f) v[f] = e == f ? r : v[f];
                        }
                    }
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int j = 32 * jt + 8 * g + 4 * h + e;
                    s1 = s1 + (j < Dm ? v[e] : 0.0f);
                }
            }
    } else {
        // SKIP: groups of eight rows past the dimension are skipped outright (a
        // uniform branch) for the objectives with transcendental terms -- worth it
        // where the row tiles are padded far past D (the wide kernel) or the terms are
        // costly (Weierstrass, Schwefel in the LDS kernel); rows j >= D are masked in
        // every case
        constexpr bool skip = SKIP && FUNC != F_SPHERE && FUNC != F_ELLIPTIC;
#pragma unroll
        for (int jt = 0; jt < NT; ++jt)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (skip && 32 * jt + 8 * (r >> 2) >= D) continue;
                const int j = 32 * jt + 8 * (r >> 2) + 4 * h + (r & 3);
                const float z = acc[jt][r];
                if constexpr (FUNC == F_SPHERE) {
                    const float v = z * z;
                    if (j < Dm) s1 = s1 + v;
                } else if constexpr (FUNC == F_ELLIPTIC) {
                    const float zz = z * z;
                    const float v = tab[j] * zz;
                    if (j < Dm) s1 = s1 + v;
                } else if constexpr (FUNC == F_ACKLEY) {
                    const float v = z * z;
                    const float sq = sinpi_sq(z);
                    if (j < Dm) { s1 = s1 + v; s2 = s2 + sq; }
                } else if constexpr (FUNC == F_RASTRIGIN) {
                    const float zz = z * z;
                    const float s20 = 20.0f * sinpi_sq(z);
                    const float v = zz + s20;
                    if (j < Dm) s1 = s1 + v;
                } else if constexpr (FUNC == F_GRIEWANK) {
                    const float v = z * z;
                    const float a = 2.0f * sinpi_sq(z * tab[j]);
                    if (j < Dm) {
                        s1 = s1 + v;
                        const float om = 1.0f - s2;
                        const float aq = a * om;
                        s2 = s2 + aq;
                    }
                } else if constexpr (FUNC == F_WEIERSTRASS) {
                    if (j < Dm) s1 = s1 + weier_term(z);
                } else if constexpr (FUNC == F_SCHWEFEL) {
                    const float sq = sqrtf(__builtin_fabsf(z));
                    const float tv = sq * 0.159154937f;
                    const float v = z * sin2pi(tv);
                    if (j < Dm) s1 = s1 + v;
                } else if constexpr (FUNC == F_REF_ACKLEY) {
                    // the reference's terms: z^2 and cospif(2 z)
                    const float v = z * z;
                    const float tz = 2.0f * z;
                    const float c = cospi_f(tz);
                    if (j < Dm) { s1 = s1 + v; s2 = s2 + c; }
                } else if constexpr (FUNC == F_REF_RASTRIGIN) {
                    // z^2 - 10 cospif(2 z) + 10
                    const float zz = z * z;
                    const float tz = 2.0f * z;
                    const float c = cospi_f(tz);
                    const float t = 10.0f * c;
                    float v = zz - t;
                    v = v + 10.0f;
                    if (j < Dm) s1 = s1 + v;
                } else if constexpr (FUNC == F_REF_GRIEWANK) {
                    // z^2 and the factor cosf(z / sqrtf(j + 1)), cos q = 1 - 2 sin^2(q / 2)
                    const float v = z * z;
                    const float q = z / tab[j];
                    const float tq = q * 0.159154937f;
                    const float t = 2.0f * sinpi_sq(tq);
                    const float c = 1.0f - t;
                    if (j < Dm) { s1 = s1 + v; s2 = s2 * c; }
                } else if constexpr (FUNC == F_REF_WEIERSTRASS) {
                    if (j < Dm) s1 = s1 + weier_ref_term(z);
                }
            }
    }
    const float o1 = __shfl_xor(s1, 32);
    const float S = s1 + o1;  // A + B (commutative)
    const float Df = (float)D;
    if constexpr (FUNC == F_ACKLEY) {
        const float o2 = __shfl_xor(s2, 32);
        const float C = s2 + o2;
        const float q = S / Df;
        const float r = sqrtf(q);
        const float a = -0.2f * r;
        const float t1 = -20.0f * expm1_f(a);
        const float c2 = C / Df;
        const float b = -2.0f * c2;
        const float t2 = 2.71828175f * expm1_f(b);
        return t1 - t2;
    } else if constexpr (FUNC == F_GRIEWANK) {
        const float o2 = __shfl_xor(s2, 32);
        const float qA = h ? o2 : s2;
        const float qB = h ? s2 : o2;
        const float om = 1.0f - qA;
        const float tq = qB * om;
        const float Qv = qA + tq;
        const float s4 = S / 4000.0f;
        return s4 + Qv;
    } else if constexpr (FUNC == F_SCHWEFEL) {
        // the constant from the task table: 418.982887 (textbook) or 418.9829 (reference)
        const float cd = tab[0] * Df;
        return cd - S;
    } else if constexpr (FUNC == F_REF_ACKLEY) {
        // -20 expf(-0.2 sqrtf(S / D)) - expf(C / D) + 20 + E with the reference's double
        // literals -0.2 and E = 2.718282 (SG-EMTF_DE.cu:913)
        const float o2 = __shfl_xor(s2, 32);
        const float C = s2 + o2;
        const float q = S / Df;
        const float r = sqrtf(q);
        const float a = (float)(-0.2 * (double)r);
        const float t1 = -20.0f * exp_f(a);
        const float c2 = C / Df;
        const float e2 = exp_f(c2);
        float u = t1 - e2;
        u = u + 20.0f;
        return (float)((double)u + 2.718282);
    } else if constexpr (FUNC == F_REF_GRIEWANK) {
        // 1 + S / 4000 - prod (the halves' partial products multiplied once)
        const float o2 = __shfl_xor(s2, 32);
        const float P = s2 * o2;
        const float s4 = S / 4000.0f;
        const float u = 1.0f + s4;
        return u - P;
    } else if constexpr (FUNC == F_REF_WEIERSTRASS) {
        // sum - D W with the reference's constant W in the task table
        const float dw = Df * tab[0];
        return S - dw;
    } else {
        return S;
    }
}

// One 32-row tile (tile jt, accumulators acc) of the objective, in two parts: the row
// terms (independent of every other tile; zn0: Rosenbrock's z of row 32 (jt + 1) for this
// column, 0 past the last tile), then their sums onto the lane's running s1, s2 -- the
// rows of tile jt in the order objective() adds them.
template <int FUNC>
__device__ __forceinline__ void tile_vals(const f16v& acc, int jt, int D, const float* tab, int h, float zn0,
                                          float (&v1)[16], float (&v2)[16]) {
    if constexpr (FUNC == F_ROSENBROCK) {
        float nb[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) nb[g] = __shfl_xor(acc[4 * g], 32);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int g = r >> 2;
            float zn;
            if ((r & 3) != 3) zn = acc[r + 1];
            else if (g < 3) zn = h ? nb[g + 1] : nb[g];
            else zn = h ? zn0 : nb[g];
            v1[r] = rosen_val(acc[r], zn);
            v2[r] = 0.0f;
        }
    } else if constexpr (FUNC == F_WEIERSTRASS) {
        // four rows at a time through the branch-free term (as in the wide kernels'
        // objective), groups of eight rows past D skipped (a uniform branch); v1 = 0 for the
        // masked rows as row_vals gives
#pragma unroll
        for (int g = 0; g < 4; ++g) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v1[4 * g + e] = 0.0f;
                v2[4 * g + e] = 0.0f;
            }
            if (32 * jt + 8 * g >= D) continue;
            float v[4];
            bool rare = false;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float a = __builtin_fabsf(acc[4 * g + e]);
                const bool lo = a >= 2.32830644e-10f, hi = a >= 8388608.0f;
                v[e] = weier_core(a, lo && !hi);
                rare = rare || (!lo && !hi && a != 0.0f);  // NaN or 0 < |z| < 2^-32
            }
            if (rare) {
#pragma unroll 1
                for (int e = 0; e < 4; ++e) {
                    float z = acc[4 * g];
                    z = e == 1 ? acc[4 * g + 1] : z;
                    z = e == 2 ? acc[4 * g + 2] : z;
                    z = e == 3 ? acc[4 * g + 3] : z;
                    const float a = __builtin_fabsf(z);
                    if (z != z || (a > 0.0f && a < 2.32830644e-10f)) {
                        const float r = weier_term(z);
#pragma unroll
                        for (int f = 0; f < 4; ++f) v[f] = e == f ? r : v[f];
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int j = 32 * jt + 8 * g + 4 * h + e;
                v1[4 * g + e] = j < D ? v[e] : 0.0f;
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            v1[r] = 0.0f;
            v2[r] = 0.0f;
            row_vals<FUNC>(acc[r], 32 * jt + 8 * (r >> 2) + 4 * h + (r & 3), D, tab, v1[r], v2[r]);
        }
    }
}

template <int FUNC>
__device__ __forceinline__ void tile_sums(const float (&v1)[16], const float (&v2)[16], int jt, int D, int h, float& s1,
                                          float& s2) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int j = 32 * jt + 8 * (r >> 2) + 4 * h + (r & 3);
        if constexpr (FUNC == F_ROSENBROCK) {
            if (j + 1 < D) s1 = s1 + v1[r];
        } else {
            row_sum<FUNC>(v1[r], v2[r], j < D, s1, s2);
        }
    }
}

// z = Q (x - o) for the 32 individuals of a tile (x: the lane's parity half)
// followed by the objective.  One MFMA k-step per gene pair (2 p, 2 p + 1), p < Dh:
// the steps past the dimension would add q * 0 products of the zero padding, which
// leave the accumulator unchanged (it is never -0), so they are skipped (a uniform
// branch, taken only in the last blocks).
template <int KB, int FUNC>
__device__ __forceinline__ float evaluate_tile(const float (&x)[4 * KB], const Geo& g, const Lds& L, int col, int h) {
    using S = Shape<KB>;
    constexpr int NT = S::NT;
    f16v acc[NT];
#pragma unroll
    for (int jt = 0; jt < NT; ++jt)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[jt][r] = 0.0f;
    const float* qrow = L.Q + col * S::RS + h * S::HS;
    const float* orow = L.o + h * S::HS;
#pragma unroll
    for (int sb = 0; sb < KB; ++sb) {
        float4 q[NT];
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) q[jt] = lds4(qrow + 32 * jt * S::RS + 4 * sb);
        const float4 o4 = lds4(orow + 4 * sb);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (sb < KB - 1 || 4 * sb + e < g.Dh) {
                const float w = x[4 * sb + e] - f4(o4, e);
#pragma unroll
                for (int jt = 0; jt < NT; ++jt)
                    acc[jt] = __builtin_amdgcn_mfma_f32_32x32x2f32(f4(q[jt], e), w, acc[jt], 0, 0, 0);
            }
        }
    }
    return objective<NT, FUNC>(acc, g.D, L.tab, h);
}

// ---------------------------------------------------------------------------
// Trial vectors
// ---------------------------------------------------------------------------
struct IdxStream {
    uint32_t w0, w1, w2, w3;
    int wi, wn, blk;
};

__device__ __forceinline__ bool idx_next(IdxStream& st, const Rng& rng, uint32_t gen, uint32_t index, uint32_t& out) {
    if (st.wi == st.wn) {
        if (st.blk == 3) return false;
        st.blk += 1;
        const U4 b = rng.block(gen, P_DE_IDX, index, 0u, (uint32_t)st.blk);
        st.w0 = b.x; st.w1 = b.y; st.w2 = b.z; st.w3 = b.w;
        st.wn = 4;
        st.wi = 0;
    }
    out = st.wi == 0 ? st.w0 : (st.wi == 1 ? st.w1 : (st.wi == 2 ? st.w2 : st.w3));
    st.wi += 1;
    return true;
}

// Rejection draw of an island-local index outside excl[0..nex) (reference
// CrossOver, SG-EMTF_DE.cu:559-573: redraw curand()%n while it collides).
__device__ __forceinline__ int draw_excl(IdxStream& st, const Rng& rng, uint32_t gen, uint32_t index, int m,
                                         const int* excl, int nex) {
    int distinct = 0;
    for (int a = 0; a < nex; ++a) {
        int dup = 0;
        for (int b = 0; b < a; ++b) dup |= (excl[b] == excl[a]);
        distinct += !dup;
    }
    uint32_t w;
    while (idx_next(st, rng, gen, index, w)) {
        const int r = (int)(w % (uint32_t)m);
        int bad = 0;
        if (distinct < m)
            for (int a = 0; a < nex; ++a) bad |= (excl[a] == r);
        if (!bad) return r;
    }
    for (int r = 0; r < m; ++r) {
        int bad = 0;
        for (int a = 0; a < nex; ++a) bad |= (excl[a] == r);
        if (!bad) return r;
    }
    return 0;
}

// DE mates of individual ie (island-local indices) and jrand: the first three words
// of its index stream usually give admissible mates (no rejection); otherwise the
// sequential rejection draw is replayed, which then consumes more words.
struct DeMates {
    int r1, r2, r3, jrand;
};

__device__ __forceinline__ DeMates de_mates(const EpochArgs& a, const Geo& g, const Rng& rng, int gen, int ie) {
    const int m = g.m, il = ie % m;
    const bool rand1 = (a.p.algo == A_DE_RAND_1_BIN);
    const U4 b0 = rng.block((uint32_t)gen, P_DE_IDX, (uint32_t)ie, 0u, 0u);
    DeMates r;
    r.jrand = (int)g.modD.mod(b0.w);
    r.r1 = (int)g.modM.mod(b0.x);
    r.r2 = (int)g.modM.mod(b0.y);
    r.r3 = (int)g.modM.mod(b0.z);
    const bool ok = rand1 ? (m >= 4 && r.r1 != il && r.r2 != il && r.r2 != r.r1 && r.r3 != il && r.r3 != r.r1 && r.r3 != r.r2)
                          : (m >= 3 && r.r2 != r.r1 && r.r3 != r.r1 && r.r3 != r.r2);
    bool replay = !ok;
    if (!ok && m >= 4) {
        // the rejection draws over the first seven words of the stream (block 0's x, y,
        // z, then block 1) in straight-line code: each word is taken by the first mate
        // still missing unless it collides with an excluded index; with m >= 4 every draw
        // has an admissible value, so this is the sequential rejection draw whenever
        // seven words suffice (a lane needing more replays it below)
        const U4 b1 = rng.block((uint32_t)gen, P_DE_IDX, (uint32_t)ie, 0u, 1u);
        const uint32_t w[7] = {b0.x, b0.y, b0.z, b1.x, b1.y, b1.z, b1.w};
        int got = 0, q1 = 0, q2 = 0, q3 = 0;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int v = (k == 0) ? r.r1 : ((k == 1) ? r.r2 : ((k == 2) ? r.r3 : (int)g.modM.mod(w[k])));
            const bool bad = (rand1 && v == il) || (got >= 1 && v == q1) || (got >= 2 && v == q2);
            const bool take = !bad && got < 3;
            q1 = (take && got == 0) ? v : q1;
            q2 = (take && got == 1) ? v : q2;
            q3 = (take && got == 2) ? v : q3;
            got += take ? 1 : 0;
        }
        r.r1 = q1;
        r.r2 = q2;
        r.r3 = q3;
        replay = got < 3;
    }
    if (replay) {
        IdxStream st{b0.x, b0.y, b0.z, 0u, 0, 3, 0};
        int ex[3];
        int nex = 0;
        if (rand1) ex[nex++] = il;
        r.r1 = draw_excl(st, rng, (uint32_t)gen, (uint32_t)ie, m, ex, nex);
        ex[nex++] = r.r1;
        r.r2 = draw_excl(st, rng, (uint32_t)gen, (uint32_t)ie, m, ex, nex);
        ex[nex++] = r.r2;
        r.r3 = draw_excl(st, rng, (uint32_t)gen, (uint32_t)ie, m, ex, nex);
    }
    return r;
}

// Genes 2 s + h, s = 4 sb .. 4 sb + 3, of a DE/x/1/bin trial: mutant vb + F (v2 - v3)
// clamped to the box, binomial crossover with the parent vi (16-bit uniforms from
// block cb: two gene blocks per Philox block).  Slots past the dimension are left as
// computed (finite, inside the box): every reader of a row takes its first D genes only,
// and the rotation meets them only in k-steps whose Q entries are zero.
__device__ __forceinline__ void de_block(float (&t)[4], const float4& vb, const float4& v2, const float4& v3,
                                         const float4& vi, const U4& cb, int sb, int h, int jrand, float F,
                                         float lo, float hi, uint32_t tcr) {
    // gene d = 2 s + h: the test d == jrand is made on 2 s against the lane's jrand - h,
    // so no per-gene index is held in a register
    const int jr = jrand - h;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int s2 = 2 * (4 * sb + e);
        const float diff = f4(v2, e) - f4(v3, e);
        const float sc = F * diff;
        float v = f4(vb, e) + sc;
        v = clamp_box(v, lo, hi);
        const uint32_t u16 = bits16(pick(cb, (uint32_t)(((sb & 1) << 1) | (e >> 1))), e & 1);
        t[e] = (u16 >= tcr && s2 != jr) ? f4(vi, e) : v;
    }
}

// Genes 2 s + h, s = 4 sb .. 4 sb + 3, of a GA child: SBX of parents va, vb
// (spread sign by the child's parity in its pair) then polynomial mutation,
// each clamped to the box, zero past the dimension.
struct GaConst {
    float lo, hi, rng, ec, em1, em, pm;
};

__device__ __forceinline__ void ga_block(float (&t)[4], const float4& va, const float4& vb, const Rng& rng, int gen,
                                         int ie, int sb, int h, bool odd, int D, const GaConst& c0) {
    const U4 sbk = rng.block((uint32_t)gen, P_GA_SBX, (uint32_t)ie, (uint32_t)h, (uint32_t)sb);
    const U4 pm0 = rng.block((uint32_t)gen, P_GA_PM, (uint32_t)ie, (uint32_t)h, (uint32_t)(2 * sb));
    const U4 pm1 = rng.block((uint32_t)gen, P_GA_PM, (uint32_t)ie, (uint32_t)h, (uint32_t)(2 * sb + 1));
    const float lo = c0.lo, hi = c0.hi, rng_ = c0.rng;
    const int dlim = D - h;  // gene 2 s + h < D  <=>  2 s < D - h
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int s2 = 2 * (4 * sb + e);
        const float p = uniform_open(pick(sbk, (uint32_t)e));
        const float tv = (p <= 0.5f) ? (2.0f * p) : (0.5f / (1.0f - p));
        float beta = pow_f(tv, c0.ec);
        if (odd) beta = -beta;
        const float t1 = 1.0f + beta;
        const float t2 = t1 * f4(va, e);
        const float t3 = 1.0f - beta;
        const float t4 = t3 * f4(vb, e);
        const float t5 = t2 + t4;
        float c = 0.5f * t5;
        c = clamp_box(c, lo, hi);
        const U4& pb = (e < 2) ? pm0 : pm1;
        const float r = uniform01(pick(pb, (uint32_t)((e & 1) * 2)));
        if (r < c0.pm) {
            const float u = uniform01(pick(pb, (uint32_t)((e & 1) * 2 + 1)));
            float delta = 1.0f, iv1, iv2, dl;
            const float num = c - lo;
            dl = num / rng_;
            iv1 = 2.0f * u;
            iv2 = 1.0f - iv1;
            if (u > 0.5f) {
                delta = -1.0f;
                const float num2 = hi - c;
                dl = num2 / rng_;
                const float omu = 1.0f - u;
                iv1 = 2.0f * omu;
                const float um = u - 0.5f;
                iv2 = 2.0f * um;
            }
            const float omd = 1.0f - dl;
            const float pw = pow_f(omd, c0.em1);
            const float a2 = iv2 * pw;
            const float a3 = iv1 + a2;
            const float a4 = pow_f(a3, c0.em);
            const float a5 = a4 - 1.0f;
            delta = delta * a5;
            const float step = delta * rng_;
            c = c + step;
            c = clamp_box(c, lo, hi);
        }
        t[e] = (8 * sb + 8 <= D || s2 < dlim) ? c : 0.0f;
    }
}

__device__ __forceinline__ GaConst ga_const(const EpochArgs& a, const TaskDev& td) {
    GaConst c;
    c.lo = td.lo;
    c.hi = td.hi;
    c.rng = td.hi - td.lo;
    c.ec = 1.0f / (a.p.eta_c + 1.0f);
    c.em1 = a.p.eta_m + 1.0f;
    c.em = 1.0f / c.em1;
    c.pm = a.p.pm;
    return c;
}

// The lane's trial half-row, block by block, each block's four genes fed straight into
// the MFMA k-steps of z = Q (x - o), so no lane holds a whole row (on gfx950 the FP32
// MFMA and the VALU share the SIMD's issue: measured, neither overlaps the other, so the
// interleaving saves registers, not time).  The same genes, k order and accumulators as
// generating the whole row first and then evaluating it.
// The trial genes are left in tr (for the selection copy); returns f of the trial.
template <int KB, int FUNC, int ALGO>
__device__ __forceinline__ float trial_eval(float (&tr)[4 * KB], const EpochArgs& a, const TaskDev& td, const Geo& g,
                                            const Lds& L, const Rng& rng, int gen, int ie, int col, int h_in) {
    using S = Shape<KB>;
    // the lane's half index is made opaque here: otherwise the compiler hoists every
    // mask and address derived from it out of the generation loop, where the lane masks
    // outgrow the scalar registers and come back from spill slots each generation
    int h = h_in;
    asm volatile("" : "+v"(h));
    constexpr int NT = S::NT;
    const int off = h * S::HS;
    const int il = ie % g.m, base = ie - il;
    f16v acc[NT];
#pragma unroll
    for (int jt = 0; jt < NT; ++jt)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[jt][r] = 0.0f;
    const float* qrow = L.Q + col * S::RS + off;
    const float* orow = L.o + off;
    // GA: the pair's mates and constants; DE: the three mates and jrand
    const float *p0, *p1, *p2, *p3;
    int jrand = 0;
    bool odd = false;
    GaConst c0{};
    if constexpr (ALGO == A_GA_SBX_PM) {
        const U4 mb = rng.block((uint32_t)gen, P_GA_MATE, (uint32_t)(ie & ~1), 0u, 0u);
        const int ra = (int)g.modM.mod(pick(mb, (uint32_t)(il & 1)));
        const int rb = (int)g.modM.mod(pick(mb, (uint32_t)((il & 1) ^ 1)));
        p0 = L.X + (base + ra) * S::RS + off;
        p1 = L.X + (base + rb) * S::RS + off;
        p2 = p3 = p0;
        c0 = ga_const(a, td);
        odd = (il & 1) != 0;
    } else {
        const bool rand1 = (a.p.algo == A_DE_RAND_1_BIN);
        const DeMates mt = de_mates(a, g, rng, gen, ie);
        p0 = L.X + (base + (rand1 ? mt.r1 : il)) * S::RS + off;
        p1 = L.X + (base + mt.r2) * S::RS + off;
        p2 = L.X + (base + mt.r3) * S::RS + off;
        p3 = L.X + ie * S::RS + off;
        jrand = mt.jrand;
    }
    U4 cb{0u, 0u, 0u, 0u};
#pragma unroll
    for (int sb = 0; sb < KB; ++sb) {
        float t[4];
        if constexpr (ALGO == A_GA_SBX_PM) {
            ga_block(t, lds4(p0 + 4 * sb), lds4(p1 + 4 * sb), rng, gen, ie, sb, h, odd, g.D, c0);
        } else {
            if ((sb & 1) == 0) cb = rng.block((uint32_t)gen, P_DE_CR, (uint32_t)ie, (uint32_t)h, (uint32_t)(sb >> 1));
            de_block(t, lds4(p0 + 4 * sb), lds4(p1 + 4 * sb), lds4(p2 + 4 * sb), lds4(p3 + 4 * sb), cb, sb, h, jrand,
                     a.p.F, td.lo, td.hi, a.p.cr_threshold);
        }
        float4 q[NT];
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) q[jt] = lds4(qrow + 32 * jt * S::RS + 4 * sb);
        const float4 o4 = lds4(orow + 4 * sb);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            tr[4 * sb + e] = t[e];
            // k-steps past the dimension add q * 0 of the zero padding: skipped
            if (sb < KB - 1 || 4 * sb + e < g.Dh) {
                const float w = t[e] - f4(o4, e);
#pragma unroll
                for (int jt = 0; jt < NT; ++jt)
                    acc[jt] = __builtin_amdgcn_mfma_f32_32x32x2f32(f4(q[jt], e), w, acc[jt], 0, 0, 0);
            }
        }
    }
    // the padded row groups are skipped for the two costliest epilogues only (measured:
    // the branch costs more than it saves for the others)
    // Q's rows past D are zero in LDS (zeroed at the epoch's start)
    return objective<NT, FUNC, FUNC == F_WEIERSTRASS || FUNC == F_SCHWEFEL || FUNC == F_REF_WEIERSTRASS, true>(
        acc, g.D, L.tab, h);
}

// Copies an n x D row-major matrix from global memory into LDS rows of stride RS in the
// [even | odd] gene layout (gene d at (d & 1) HS + d / 2), eight loads in flight per
// thread, so a load's latency is paid once per eight elements rather than per element.
template <int RS, int HS>
__device__ __forceinline__ void stage_rows(float* dst, const float* __restrict__ src, int n, int D, int tid) {
    const int total = n * D;
    for (int e0 = tid; e0 < total; e0 += 8 * kThreads) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int e = e0 + k * kThreads;
            v[k] = e < total ? src[e] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int e = e0 + k * kThreads;
            if (e < total) {
                const int j = e / D, d = e - j * D;
                dst[j * RS + (d & 1) * HS + (d >> 1)] = v[k];
            }
        }
    }
}

template <int KB>
__device__ __forceinline__ void load_half_row(float (&x)[4 * KB], const float* row) {
#pragma unroll
    for (int sb = 0; sb < KB; ++sb) {
        const float4 v = lds4(row + 4 * sb);
        x[4 * sb] = v.x; x[4 * sb + 1] = v.y; x[4 * sb + 2] = v.z; x[4 * sb + 3] = v.w;
    }
}

template <int KB>
__device__ __forceinline__ void store_half_row(float* row, const float (&x)[4 * KB]) {
#pragma unroll
    for (int sb = 0; sb < KB; ++sb)
        *reinterpret_cast<float4*>(row + 4 * sb) = make_float4(x[4 * sb], x[4 * sb + 1], x[4 * sb + 2], x[4 * sb + 3]);
}

// ---------------------------------------------------------------------------
// Population-level phases (all threads of the workgroup participate)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float key_of(float f) { return (f != f) ? __builtin_inff() : f; }

// ord[q m + rank] = i: the members of each island q ranked by (fitness, index)
template <int NTHR = kThreads>
__device__ __forceinline__ void rank_islands(const Lds& L, int pop, int m, int tid) {
    for (int i = tid; i < pop; i += NTHR) {
        const int q = i / m, base = q * m, il = i - base;
        const float fi = key_of(L.fit[i]);
        int r = 0;
        for (int p = 0; p < m; ++p) {
            const float fp = key_of(L.fit[base + p]);
            r += (fp < fi || (fp == fi && p < il)) ? 1 : 0;
        }
        L.ord[base + r] = i;
    }
}

template <int KB, int FUNC, int ALGO>
__device__ __forceinline__ void generation_step(const EpochArgs& a, const TaskDev& td, const Geo& g, const Lds& L,
                                                const Rng& rng, int gen, int tid) {
    using S = Shape<KB>;
    const int wave = tid >> 6, lane = tid & 63, col = lane & 31, h = lane >> 5;
    const bool multi = g.ntiles > kThreads / 64;
    float tr[4 * KB];
    float fchild = 0.0f;
    int i_keep = -1;
    bool acc_keep = false;
    for (int it = wave; it < g.ntiles; it += kThreads / 64) {
        const int i = it * 32 + col;
        const bool valid = i < g.pop;
        const int ie = valid ? i : 0;
        const float f = trial_eval<KB, FUNC, ALGO>(tr, a, td, g, L, rng, gen, ie, col, h);
        const bool accept = valid && (L.fit[i] >= f);
        if (multi) {
            if (accept) store_half_row<KB>(a.scratch + td.scr_off + (int64_t)i * S::RS + h * S::HS, tr);
            if (valid && h == 0) { L.aux[i] = accept ? 1 : 0; L.ftmp[i] = f; }
        } else {
            i_keep = i;
            acc_keep = accept;
            fchild = f;
        }
    }
    __syncthreads();
    if (!multi) {
        if (i_keep >= 0 && acc_keep) {
            store_half_row<KB>(L.X + i_keep * S::RS + h * S::HS, tr);
            if (h == 0) L.fit[i_keep] = fchild;
        }
    } else {
        const float* src = a.scratch + td.scr_off;
        for (int e = tid; e < g.pop * (2 * KB); e += kThreads) {
            const int row = e / (2 * KB), blk = e - row * (2 * KB);
            if (L.aux[row]) {
                const int c = (blk < KB) ? 4 * blk : S::HS + 4 * (blk - KB);
                *reinterpret_cast<float4*>(L.X + row * S::RS + c) =
                    *reinterpret_cast<const float4*>(src + (int64_t)row * S::RS + c);
            }
        }
        for (int i = tid; i < g.pop; i += kThreads)
            if (L.aux[i]) L.fit[i] = L.ftmp[i];
    }
    __syncthreads();
}

// Cross-island resampling (reference SharedToGlobal + SelectFromBlocks).
template <int KB, int NTHR = kThreads>
__device__ __forceinline__ void island_exchange(const EpochArgs& a, const Geo& g, const Lds& L, float* scr,
                                                const Rng& rng, int gen, int tid) {
    constexpr int RS = Shape<KB>::RS;
    const int n = a.p.n_islands, m = g.m;
    for (int q = tid; q < n; q += NTHR) {
        int* arr = L.aux + q * m;
        for (int p = 0; p < m; ++p) arr[p] = p;
        for (int k = m - 1; k >= 0; --k) {
            const uint32_t w = rng.word((uint32_t)gen, P_ISLX, (uint32_t)q, 0u, (uint32_t)(m - 1 - k));
            const int t = (int)(w % (uint32_t)(k + 1));
            const int tmp = arr[k];
            arr[k] = arr[t];
            arr[t] = tmp;
        }
        for (int p = 0; p < m; ++p) L.ord[q * m + arr[p]] = p;  // inverse permutation
    }
    __syncthreads();
    // one island offset per slot k, shared by every island (the reference's
    // select_interval[iter * IPB + k] has no block term, SG-EMTF.cu:1227, :1597): each slot
    // is a rotation of the islands, so the exchange permutes the task's rows
    for (int i = tid; i < g.pop; i += NTHR) {
        const int q = i / m, k = i - q * m;
        const uint32_t w = rng.word((uint32_t)gen, P_ISLX, 0u, 1u, (uint32_t)k);
        const int src = (q + (int)(w % (uint32_t)n)) % n;
        L.aux[i] = src * m + L.ord[src * m + k];
    }
    __syncthreads();
    for (int e = tid; e < g.pop * RS; e += NTHR) {
        const int row = e / RS, c = e - row * RS;
        scr[e] = L.X[L.aux[row] * RS + c];
    }
    for (int i = tid; i < g.pop; i += NTHR) L.ftmp[i] = L.fit[L.aux[i]];
    __syncthreads();
    for (int e = tid; e < g.pop * RS; e += NTHR) L.X[e] = scr[e];
    for (int i = tid; i < g.pop; i += NTHR) L.fit[i] = L.ftmp[i];
    __syncthreads();
}

// Intra-task island migration (reference IslandMigration).
template <int KB, int NTHR = kThreads>
__device__ __forceinline__ void island_migration(const EpochArgs& a, const Geo& g, const Lds& L, int tid) {
    constexpr int RS = Shape<KB>::RS;
    const int n = a.p.n_islands, m = g.m, kk = a.p.migrate_count;
    rank_islands<NTHR>(L, g.pop, m, tid);
    __syncthreads();
    for (int e = tid; e < n * kk; e += NTHR) {
        const int q = e / kk, j = e - q * kk, p = (q + 1) % n;
        const int src = L.ord[q * m + j], dst = L.ord[p * m + m - kk + j];
        L.aux[e] = (L.fit[src] < L.fit[dst]) ? 1 : 0;
    }
    __syncthreads();
    for (int e = tid; e < n * kk * RS; e += NTHR) {
        const int pr = e / RS, c = e - pr * RS;
        if (L.aux[pr]) {
            const int q = pr / kk, j = pr - q * kk, p = (q + 1) % n;
            const int src = L.ord[q * m + j], dst = L.ord[p * m + m - kk + j];
            L.X[dst * RS + c] = L.X[src * RS + c];
        }
    }
    for (int e = tid; e < n * kk; e += NTHR) {
        if (L.aux[e]) {
            const int q = e / kk, j = e - q * kk, p = (q + 1) % n;
            const int src = L.ord[q * m + j], dst = L.ord[p * m + m - kk + j];
            L.fit[dst] = L.fit[src];
        }
    }
    __syncthreads();
}

// Inter-task transfer at the start of an epoch (reference knowledgeTransfer).
template <int KB, int FUNC>
__device__ __forceinline__ void transfer_in(const EpochArgs& a, const TaskDev& td, const Geo& g, const Lds& L,
                                            const Rng& rng, int tid, int eg0) {
    using S = Shape<KB>;
    // draws at the task's own generation g0, the mailbox buffer by the handle's clock eg0
    const int k = a.p.transfer_count, TI = a.p.transfer_interval, g0 = eg0 - td.gen_base;
    const int wave = tid >> 6, lane = tid & 63, col = lane & 31, h = lane >> 5;
    rank_islands(L, g.pop, g.pop, tid);
    __syncthreads();
    const uint32_t wd = rng.word((uint32_t)g0, P_MIG, 0u, 0u, 0u);
    const int dn = pick_donor(a, (int)rng.task, wd);
    const TaskDev dd = a.tasks[dn];
    const float* mb = a.mailbox + (int64_t)((eg0 / TI) & 1) * a.mailbox_stride + dd.mb_off;
    const bool same_box = (dd.lo == td.lo && dd.hi == td.hi);
    const float rng_s = td.hi - td.lo, rng_d = dd.hi - dd.lo;
    const int ntc = (k + 31) / 32;
    for (int ct = wave; ct < ntc; ct += kThreads / 64) {
        const int j = ct * 32 + col;
        const bool valid = j < k;
        const int je = valid ? j : 0;
        float cand[S::HS];
        load_half_row<KB>(cand, L.X + L.ord[je] * S::RS + h * S::HS);
#pragma unroll
        for (int s = 0; s < S::HS; ++s) {
            const int d = 2 * s + h;
            if (d < g.D && d < dd.dim) {
                bool take = true;
                if (a.p.transfer_cr > 0.0f) {
                    const uint32_t wu = rng.word((uint32_t)g0, P_XFER, (uint32_t)je, (uint32_t)h, (uint32_t)s);
                    take = uniform01(wu) > a.p.transfer_cr;
                }
                if (take) {
                    const float v = mb[(int64_t)je * dd.dim + d];
                    if (same_box) cand[s] = v;
                    else {
                        const float t1 = v - dd.lo;
                        const float t2 = t1 / rng_d;
                        const float t3 = t2 * rng_s;
                        cand[s] = td.lo + t3;
                    }
                }
            }
        }
        const float f = evaluate_tile<KB, FUNC>(cand, g, L, col, h);
        if (valid) {
            const int tgt = L.ord[g.pop - k + j];
            if (f < L.fit[tgt]) {
                store_half_row<KB>(L.X + tgt * S::RS + h * S::HS, cand);
                if (h == 0) L.fit[tgt] = f;
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Stand-alone objective and adjoint (emtf_objective_fwd / emtf_objective_adjoint).
// A wave evaluates 32 points.  The forward pass is the engine's evaluator (bit
// for bit).  The adjoint forms dz = df/dz per dimension in binary32 from z
// staged in LDS, scales it by the upstream gradient, and contracts
// gx = Q^T dz with a second run of v_mfma_f32_32x32x2_f32 (A = Q read
// transposed from LDS, B = dz with one point per column).
// ---------------------------------------------------------------------------
template <int KB>
struct ObjShape {
    static constexpr int PB = (KB <= 8) ? 128 : 32;  // points per workgroup
    static constexpr int ZS = 32 * Shape<KB>::NT + 1; // LDS stride of staged z rows
};

template <int KB>
__host__ __device__ inline int obj_lds_floats() {
    using S = Shape<KB>;
    constexpr int PB = ObjShape<KB>::PB, ZS = ObjShape<KB>::ZS;
    return 32 * S::NT * S::RS + S::RS + 32 * S::NT + 2 * PB * ZS;
}

// Per-point factors of d f / d z: Ackley's two (from the sums over the point's z).
struct DzConst {
    float ka, kb;
};

// Ackley's gradient factors from its sums s1 = sum z_j^2, s2 = sum sin^2(pi z_j).
__device__ __forceinline__ DzConst ackley_const(float s1, float s2, int D) {
    DzConst k;
    const float Df = (float)D;
    const float r = sqrtf(s1 / Df);
    const float e1 = expm1_f(-0.2f * r) + 1.0f;
    const float e2 = expm1_f(-2.0f * (s2 / Df)) + 1.0f;
    k.ka = (r > 0.0f) ? (4.0f * e1) / (Df * r) : 0.0f;
    k.kb = (6.28318548f * 2.71828175f / Df) * e2;
    return k;
}

template <int FUNC>
__device__ __forceinline__ DzConst dz_const(const float* z, int D) {
    DzConst k{0.0f, 0.0f};
    if constexpr (FUNC == F_ACKLEY) {
        float s1 = 0.0f, s2 = 0.0f;
        for (int j = 0; j < D; ++j) { s1 = s1 + z[j] * z[j]; s2 = s2 + sinpi_sq(z[j]); }
        k = ackley_const(s1, s2, D);
    }
    return k;
}

// Griewank's factor cos(theta_j) = 1 - 2 sin^2(theta_j / 2), theta_j / (2 pi) = z_j tab[j]
__device__ __forceinline__ float griewank_cos(float zj, float tj) { return 1.0f - 2.0f * sinpi_sq(zj * tj); }

// d f / d z_j (0 for j >= D) of the point z; Griewank takes pre = prod_{k<j} cos(theta_k)
// and suf = prod_{k>j} cos(theta_k), accumulated from the top.
template <int FUNC>
__device__ __forceinline__ float dz_term(int j, const float* z, int D, const float* tab, const DzConst& kc, float pre,
                                         float suf) {
    float v = 0.0f;
    if (j < D) {
        const float zj = z[j];
        if constexpr (FUNC == F_SPHERE) v = 2.0f * zj;
        else if constexpr (FUNC == F_ELLIPTIC) v = 2.0f * tab[j] * zj;
        else if constexpr (FUNC == F_RASTRIGIN) v = 2.0f * zj + 62.8318531f * sin2pi(zj);
        else if constexpr (FUNC == F_ACKLEY) v = kc.ka * zj + kc.kb * sin2pi(zj);
        else if constexpr (FUNC == F_ROSENBROCK) {
            if (j + 1 < D) v = 400.0f * zj * (zj * zj - z[j + 1]) + 2.0f * (zj - 1.0f);
            if (j > 0) v = v - 200.0f * (z[j - 1] * z[j - 1] - zj);
        } else if constexpr (FUNC == F_GRIEWANK) {
            const float sn = sin2pi(zj * tab[j]);  // sin(theta_j)
            v = zj / 2000.0f + pre * suf * sn * (6.28318548f * tab[j]);
        } else if constexpr (FUNC == F_WEIERSTRASS) {
            // sum_k 2^(1-k) pi 3^k sin(2 pi 3^k z), fractions of 3^k |z| as in the forward pass
            const float a = __builtin_fabsf(zj);
            float acc = 0.0f;
            if (a < 8388608.0f) {
                float c = 6.28318548f;  // 2 pi (1.5)^k
                if (a < 2.32830644e-10f) {
                    float p3 = 1.0f;
                    for (int k = 0; k <= 20; ++k) {
                        acc = fmaf(c, sin2pi(a * p3), acc);
                        c = c * 1.5f;
                        p3 = p3 * 3.0f;
                    }
                } else {
                    const uint32_t u = __float_as_uint(a);
                    const int ex = (int)((u >> 23) & 0xFFu) - 127;
                    const uint64_t m = (uint64_t)((u & 0x007FFFFFu) | 0x00800000u);
                    uint64_t y = m << (64 - (23 - ex));
                    for (int k = 0; k <= 20; ++k) {
                        const float r = (float)(uint32_t)(y >> 32) * 2.3283064365386963e-10f;
                        acc = fmaf(c, sin2pi(r), acc);
                        c = c * 1.5f;
                        y = y + (y << 1);
                    }
                }
            }
            v = __builtin_copysignf(1.0f, zj) * acc;
        } else if constexpr (FUNC == F_SCHWEFEL) {
            const float sq = sqrtf(__builtin_fabsf(zj));
            const float t = sq * 0.159154937f;
            v = -(sin2pi(t) + 0.5f * sq * sin2pi(t + 0.25f));
        }
    }
    return v;
}

// d f / d z_j of the lane's point for j = 2 s + h (the B-operand layout of the
// second contraction), from the full z row in LDS.
template <int KB, int FUNC>
__device__ __forceinline__ void dfdz(float (&g)[4 * KB], const float* z, float* pre, int D, const float* tab, int h) {
    constexpr int HS = 4 * KB;
    if constexpr (FUNC == F_GRIEWANK) {
        float p = 1.0f;  // prefix products
        for (int j = 0; j < D; ++j) {
            pre[j] = p;
            p = p * griewank_cos(z[j], tab[j]);
        }
    }
    const DzConst kc = dz_const<FUNC>(z, D);
    float suf = 1.0f;  // Griewank suffix product, walked downwards
    int jn = D - 1;
#pragma unroll
    for (int s = HS - 1; s >= 0; --s) {
        const int j = 2 * s + h;
        if constexpr (FUNC == F_GRIEWANK) {
            while (jn > j) {
                suf = suf * griewank_cos(z[jn], tab[jn]);
                --jn;
            }
        }
        g[s] = dz_term<FUNC>(j, z, D, tab, kc, (FUNC == F_GRIEWANK && j < D) ? pre[j] : 0.0f, suf);
    }
}

// z rows of tile jt for the lane's column from w = x - o (the lane's parity half, in
// registers): the k-steps of evaluate_tile's chain for that tile, whole blocks of four
// from the first, stopping after the block that holds step Dh - 1 (the steps past the
// dimension in it add zeros).
template <int KB>
__device__ __forceinline__ f16v tile_z(const float (&w)[4 * KB], const float* Q, int jt, int Dh, int col, int h) {
    using S = Shape<KB>;
    f16v acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const float* qrow = Q + (32 * jt + col) * S::RS + h * S::HS;
#pragma unroll
    for (int sb = 0; sb < KB; ++sb) {
        if (4 * sb >= Dh) break;
        const float4 q = lds4(qrow + 4 * sb);
#pragma unroll
        for (int e = 0; e < 4; ++e) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(f4(q, e), w[4 * sb + e], acc, 0, 0, 0);
    }
    return acc;
}

// Polish staging per elite slot (floats): the point in the [even | odd] row layout, z,
// dz, the gradient and Griewank's prefix / suffix products (natural order), scalars,
// and the line search's candidates of one round (fitness, step) -- cpe each.
template <int KB>
struct PolishSlot {
    static constexpr int RS = Shape<KB>::RS, ZS = 32 * Shape<KB>::NT + 4;  // rows 16-byte aligned
    __host__ __device__ static constexpr int floats(int cpe) { return (RS + 5 * ZS + 16 + 2 * cpe + 3) & ~3; }
    static constexpr int kScratch = 128 + 128;  // after the slots: line-search sums, Rosenbrock rows
};

// Armijo polish of the top-k elites of one task, run between epochs after generation g1
// (restates the oracle's polish_task): from a Polyak step alpha = f / |g|^2 along -g,
// halve alpha until f(clamp(x - alpha g)) <= f - 1e-4 alpha |g|^2 (at most 24 halvings),
// take the step, repeat.  The elites go through in groups of 32 / cpe slots, all four
// waves working on every step of a group: each wave computes one 32-row tile of
// z = Q (x - o) (first step only) and of g = Q^T dz for every slot (column c holds slot
// c / cpe); dz is spread over all threads by (slot, dimension); the line search evaluates
// cpe halvings per slot per round (column c: halving round base + c mod cpe), each wave
// one row tile of the candidates' z, and the smallest accepted halving wins -- the step
// the one-at-a-time search takes; its z is kept as the next step's.  At a transfer
// boundary the task's elites are published afterwards.
template <int KB, int FUNC>
__device__ __forceinline__ void polish_body(const EpochArgs& a, float* smem, const TaskDev& td) {
    using S = Shape<KB>;
    constexpr int RS = S::RS, HS = S::HS, NT = S::NT, ZS = PolishSlot<KB>::ZS;
    constexpr int NW = kThreads / 64;
    const int tid = threadIdx.x, lane = tid & 63, col = lane & 31, h = lane >> 5, wave = tid >> 6;
    Geo g{};
    g.D = td.dim;
    g.Dh = (g.D + 1) >> 1;
    g.pop = td.pop;
    g.popP = (g.pop + 31) & ~31;
    Lds L{};
    L.Q = smem;
    L.o = L.Q + 32 * NT * RS;
    L.tab = L.o + RS;
    L.fit = L.tab + 32 * NT;
    L.ord = reinterpret_cast<int*>(L.fit + g.popP);
    const int cpe = a.polish_cpe, nsl = 32 / cpe, PSZ = PolishSlot<KB>::floats(cpe);
    float* const slots = reinterpret_cast<float*>(L.ord + g.popP);
    float* const PS = slots + nsl * PSZ;  // the line search's running row sums, s1 | s2 per lane
    float* const PZ = PS + 128;           // Rosenbrock: row 0 of each tile per candidate
    static_assert(NT <= NW, "one row tile per wave in the line search");
    const int nz = 32 * NT * RS + RS + 32 * NT;
    for (int e = tid; e < nz; e += kThreads) smem[e] = 0.0f;
    __syncthreads();
    stage_rows<RS, HS>(L.Q, a.Q + td.q_off, g.D, g.D, tid);
    for (int d = tid; d < g.D; d += kThreads) {
        L.o[(d & 1) * HS + (d >> 1)] = a.shift[td.v_off + d];
        L.tab[d] = a.tab[td.v_off + d];
    }
    float* fitg = a.fit + td.fit_off;
    float* Xg = a.X + td.x_off;
    for (int i = tid; i < g.pop; i += kThreads) L.fit[i] = fitg[i];
    __syncthreads();
    rank_islands(L, g.pop, g.pop, tid);
    __syncthreads();
    const int K = a.p.polish_topk < g.pop ? a.p.polish_topk : g.pop;
    // per-slot views
    auto X = [&](int sl) { return slots + sl * PSZ; };
    auto Z = [&](int sl) { return X(sl) + RS; };
    auto DZ = [&](int sl) { return Z(sl) + ZS; };
    auto G = [&](int sl) { return DZ(sl) + ZS; };
    auto PRE = [&](int sl) { return G(sl) + ZS; };
    auto SUF = [&](int sl) { return PRE(sl) + ZS; };
    auto SF = [&](int sl) { return SUF(sl) + ZS; };  // f, g2, alpha0, ka, kb, fnew, alpha_ok
    auto SI = [&](int sl) { return reinterpret_cast<int*>(SF(sl) + 8); };  // active, done, first, elite
    auto CF = [&](int sl) { return SF(sl) + 16; };
    auto CA = [&](int sl) { return CF(sl) + cpe; };
    constexpr int kNone = 1 << 30;
    const int my_sl = col / cpe, sub = col - my_sl * cpe;
    for (int gb = 0; gb < K; gb += nsl) {
        const int ns = (K - gb) < nsl ? (K - gb) : nsl;
        for (int e = tid; e < nsl * 2 * HS; e += kThreads) {
            const int sl = e / (2 * HS), c = e - sl * (2 * HS);
            const int d = (c < HS) ? 2 * c : 2 * (c - HS) + 1;
            float v = 0.0f;
            if (sl < ns && d < g.D) v = Xg[(int64_t)L.ord[gb + sl] * g.D + d];
            X(sl)[c] = v;
        }
        for (int sl = tid; sl < nsl; sl += kThreads) {
            SF(sl)[0] = sl < ns ? L.fit[L.ord[gb + sl]] : 0.0f;
            SI(sl)[0] = sl < ns ? 1 : 0;
            SI(sl)[3] = sl < ns ? L.ord[gb + sl] : 0;
        }
        __syncthreads();
        for (int step = 0; step < a.p.polish_steps; ++step) {
            // z = Q (x - o) of every slot, one 32-row tile per wave (after the first step the
            // line search left it: an accepted slot's z, and an inactive slot's is unused)
            for (int jt = wave; step == 0 && jt < NT; jt += NW) {
                const float* xr = X(my_sl) + h * HS;
                float w[HS];
#pragma unroll
                for (int s4 = 0; s4 < HS; s4 += 4) {
                    const float4 x4 = lds4(xr + s4), o4 = lds4(L.o + h * HS + s4);
#pragma unroll
                    for (int e = 0; e < 4; ++e) w[s4 + e] = f4(x4, e) - f4(o4, e);
                }
                const f16v acc = tile_z<KB>(w, L.Q, jt, g.Dh, col, h);
                if (sub == 0) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) Z(my_sl)[32 * jt + 8 * (r >> 2) + 4 * h + (r & 3)] = acc[r];
                }
            }
            __syncthreads();
            // per-slot factors: Ackley's, Griewank's prefix (even threads) / suffix (odd)
            // products; their per-dimension terms first, over all (slot, dimension) pairs
            // (into the dz and prefix rows, free until the dz pass), then the sums and
            // products in dimension order, one thread per slot and direction
            if constexpr (FUNC == F_ACKLEY || FUNC == F_GRIEWANK) {
                for (int e = tid; e < ns * g.D; e += kThreads) {
                    const int sl = e / g.D, j = e - sl * g.D;
                    const float zj = Z(sl)[j];
                    if constexpr (FUNC == F_ACKLEY) {
                        DZ(sl)[j] = zj * zj;
                        PRE(sl)[j] = sinpi_sq(zj);
                    } else {
                        DZ(sl)[j] = griewank_cos(zj, L.tab[j]);
                    }
                }
                __syncthreads();
            }
            if constexpr (FUNC == F_ACKLEY) {
                for (int sl = tid; sl < ns; sl += kThreads) {
                    float s1 = 0.0f, s2 = 0.0f;
                    for (int j = 0; j < g.D; ++j) {
                        s1 = s1 + DZ(sl)[j];
                        s2 = s2 + PRE(sl)[j];
                    }
                    const DzConst kc = ackley_const(s1, s2, g.D);
                    SF(sl)[3] = kc.ka;
                    SF(sl)[4] = kc.kb;
                }
                __syncthreads();
            }
            if constexpr (FUNC == F_GRIEWANK) {
                for (int e = tid; e < 2 * ns; e += kThreads) {
                    const int sl = e >> 1;
                    const float* c = DZ(sl);
                    float p = 1.0f;
                    if ((e & 1) == 0) {
                        for (int j = 0; j < g.D; ++j) {
                            PRE(sl)[j] = p;
                            p = p * c[j];
                        }
                    } else {
                        for (int j = g.D - 1; j >= 0; --j) {
                            SUF(sl)[j] = p;
                            p = p * c[j];
                        }
                    }
                }
                __syncthreads();
            }
            // dz of every (slot, dimension); zero past the dimension and for empty slots
            for (int e = tid; e < nsl * 2 * HS; e += kThreads) {
                const int sl = e / (2 * HS), j = e - sl * (2 * HS);
                float v = 0.0f;
                if (sl < ns) {
                    const DzConst kc{SF(sl)[3], SF(sl)[4]};
                    const bool gw = FUNC == F_GRIEWANK && j < g.D;
                    v = dz_term<FUNC>(j, Z(sl), g.D, L.tab, kc, gw ? PRE(sl)[j] : 0.0f, gw ? SUF(sl)[j] : 0.0f);
                }
                DZ(sl)[j] = v;
            }
            __syncthreads();
            // g = Q^T dz of every slot, one 32-row tile of dimensions per wave
            for (int dt = wave; dt < NT; dt += NW) {
                const float* dzr = DZ(my_sl);
                const int d = 32 * dt + col;
                f16v gacc;
#pragma unroll
                for (int r = 0; r < 16; ++r) gacc[r] = 0.0f;
                // blocks of four k-steps (rows j = 2 s2 + h); the steps past the dimension
                // in the last block add zeros (Q's padding rows and dz past D are zero)
                const float* qc = L.Q + (d & 1) * HS + (d >> 1);
#pragma unroll
                for (int sb = 0; sb < KB; ++sb) {
                    if (4 * sb >= g.Dh) break;
                    float qv[4], zv[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int j = 2 * (4 * sb + e) + h;
                        qv[e] = qc[j * RS];
                        zv[e] = dzr[j];
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) gacc = __builtin_amdgcn_mfma_f32_32x32x2f32(qv[e], zv[e], gacc, 0, 0, 0);
                }
                if (sub == 0) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) G(my_sl)[32 * dt + 8 * (r >> 2) + 4 * h + (r & 3)] = gacc[r];
                }
            }
            __syncthreads();
            // |g|^2 summed in dimension order, the Polyak step
            for (int sl = tid; sl < ns; sl += kThreads) {
                const float* gr = G(sl);
                float g2 = 0.0f;
                int d = 0;
                for (; d + 8 <= g.D; d += 8) {  // eight loads in flight, the chain in order
                    float v[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) v[k] = gr[d + k];
#pragma unroll
                    for (int k = 0; k < 8; ++k) g2 = fmaf(v[k], v[k], g2);
                }
                for (; d < g.D; ++d) g2 = fmaf(gr[d], gr[d], g2);
                const float f = SF(sl)[0];
                SF(sl)[1] = g2;
                SF(sl)[2] = f / g2;
                if (!(g2 > 0.0f) || !(f > 0.0f)) SI(sl)[0] = 0;
                SI(sl)[1] = 0;
                SI(sl)[2] = kNone;
            }
            __syncthreads();
            // Armijo rounds: halvings r0 .. r0 + cpe - 1 of every slot (column col: slot
            // col / cpe, halving r0 + col % cpe).  Wave w computes row tile w of the 32
            // candidates' z, and the objective's row sums run through the tiles in order:
            // wave w adds tile w's rows onto the sums wave w - 1 left in LDS, the last wave
            // finishes the fitness.  The accepted candidate's z is the next step's z.
            for (int r0 = 0;; r0 += cpe) {
                const int b = r0 + sub;
                const bool act = my_sl < ns && SI(my_sl)[0] && !SI(my_sl)[1];
                float alpha = SF(my_sl)[2];
                for (int q = 0; q < b; ++q) alpha = alpha * 0.5f;
                const float* xr = X(my_sl) + h * HS;
                const float* gr = G(my_sl);
                f16v acc;
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
                if (wave < NT) {
                    float w[HS];
#pragma unroll
                    for (int s4 = 0; s4 < HS; s4 += 4) {
                        const float4 x4 = lds4(xr + s4), o4 = lds4(L.o + h * HS + s4);
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            const int s2 = s4 + e;
                            const float st = alpha * gr[2 * s2 + h];  // in the row (zeros past D)
                            const float xc = clampf(f4(x4, e) - st, td.lo, td.hi);
                            const float xn = (2 * s2 + h < g.D) ? xc : 0.0f;
                            w[s2] = xn - f4(o4, e);
                        }
                    }
                    acc = tile_z<KB>(w, L.Q, wave, g.Dh, col, h);
                    if (FUNC == F_ROSENBROCK && h == 0) PZ[32 * wave + col] = acc[0];
                }
                __syncthreads();
                float v1[16], v2[16];
                if (wave < NT) {
                    const float zn0 = (FUNC == F_ROSENBROCK && wave + 1 < NT) ? PZ[32 * (wave + 1) + col] : 0.0f;
                    tile_vals<FUNC>(acc, wave, g.D, L.tab, h, zn0, v1, v2);
                }
                float s1 = 0.0f, s2 = 0.0f;
                for (int st = 0; st < NT; ++st) {
                    if (wave == st) {
                        if (st > 0) {
                            s1 = PS[lane];
                            s2 = PS[64 + lane];
                        }
                        tile_sums<FUNC>(v1, v2, st, g.D, h, s1, s2);
                        if (st + 1 < NT) {
                            PS[lane] = s1;
                            PS[64 + lane] = s2;
                        }
                    }
                    if (st + 1 < NT) __syncthreads();
                }
                if (wave == NT - 1) {
                    const float fn = objective_tail<FUNC>(s1, s2, g.D, L.tab, h);
                    const float c1 = 1.0e-4f * alpha;
                    const float c2 = c1 * SF(my_sl)[1];
                    const float thr = SF(my_sl)[0] - c2;
                    if (act && b < 24 && fn <= thr && h == 0) {
                        CF(my_sl)[sub] = fn;
                        CA(my_sl)[sub] = alpha;
                        atomicMin(&SI(my_sl)[2], b);
                    }
                }
                __syncthreads();
                for (int sl = tid; sl < ns; sl += kThreads) {
                    const int fb = SI(sl)[2];
                    if (SI(sl)[0] && !SI(sl)[1] && fb != kNone) {
                        SI(sl)[1] = 1;
                        SF(sl)[5] = CF(sl)[fb - r0];
                        SF(sl)[6] = CA(sl)[fb - r0];
                    }
                }
                __syncthreads();
                // the accepted halving's lanes keep z = Q (x_new - o) for the next step (the
                // same MFMA chain the forward pass would run on the stepped point)
                if (wave < NT && my_sl < ns && SI(my_sl)[2] == b) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) Z(my_sl)[32 * wave + 8 * (r >> 2) + 4 * h + (r & 3)] = acc[r];
                }
                bool more = false;
                for (int sl = 0; sl < ns; ++sl) more = more || (SI(sl)[0] && !SI(sl)[1]);
                if (!more || r0 + cpe >= 24) break;
            }
            // accepted slots step to clamp(x - alpha g), the others stop
            for (int e = tid; e < ns * 2 * HS; e += kThreads) {
                const int sl = e / (2 * HS), c = e - sl * (2 * HS);
                if (SI(sl)[0] && SI(sl)[1]) {
                    const int d = (c < HS) ? 2 * c : 2 * (c - HS) + 1;
                    const float st = SF(sl)[6] * G(sl)[d];
                    X(sl)[c] = (d < g.D) ? clampf(X(sl)[c] - st, td.lo, td.hi) : 0.0f;
                }
            }
            __syncthreads();
            for (int sl = tid; sl < ns; sl += kThreads) {
                if (SI(sl)[0] && SI(sl)[1]) SF(sl)[0] = SF(sl)[5];
                else SI(sl)[0] = 0;
            }
            __syncthreads();
        }
        for (int e = tid; e < ns * g.D; e += kThreads) {
            const int sl = e / g.D, d = e - sl * g.D;
            Xg[(int64_t)SI(sl)[3] * g.D + d] = X(sl)[(d & 1) * HS + (d >> 1)];
        }
        for (int sl = tid; sl < ns; sl += kThreads) {
            L.fit[SI(sl)[3]] = SF(sl)[0];
            fitg[SI(sl)[3]] = SF(sl)[0];
        }
        __syncthreads();
    }
    if (a.publish) {
        rank_islands(L, g.pop, g.pop, tid);
        __syncthreads();
        const int k = a.p.transfer_count, TI = a.p.transfer_interval;
        float* mb = a.mailbox + (int64_t)((a.g1 / TI) & 1) * a.mailbox_stride + td.mb_off;
        for (int e = tid; e < k * g.D; e += kThreads) {
            const int r = e / g.D, d = e - r * g.D;
            mb[e] = Xg[(int64_t)L.ord[r] * g.D + d];
        }
    }
}

template <int KB>
__global__ __launch_bounds__(kThreads) void polish_kernel(EpochArgs a, const int* __restrict__ order) {
    extern __shared__ float smem[];
    const TaskDev td = a.tasks[order[blockIdx.x]];
    switch (td.func) {
    case F_SPHERE: polish_body<KB, F_SPHERE>(a, smem, td); break;
    case F_ROSENBROCK: polish_body<KB, F_ROSENBROCK>(a, smem, td); break;
    case F_ACKLEY: polish_body<KB, F_ACKLEY>(a, smem, td); break;
    case F_RASTRIGIN: polish_body<KB, F_RASTRIGIN>(a, smem, td); break;
    case F_GRIEWANK: polish_body<KB, F_GRIEWANK>(a, smem, td); break;
    case F_WEIERSTRASS: polish_body<KB, F_WEIERSTRASS>(a, smem, td); break;
    case F_SCHWEFEL: polish_body<KB, F_SCHWEFEL>(a, smem, td); break;
    case F_ELLIPTIC: polish_body<KB, F_ELLIPTIC>(a, smem, td); break;
    default: break;
    }
}

// Thread 0's bounded wait for *p >= v under the persistent schedule (device-scope acquire
// loads; false once the handle's error flag is set or the wait runs out, which sets it).
__device__ __noinline__ bool persist_wait(const int32_t* p, int32_t v, const PersistArgs& pa) {
    const uint64_t t0 = wall_clock64();
    for (;;) {
        if (__hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= v) return true;
        if (__hip_atomic_load(pa.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
        if ((int64_t)(wall_clock64() - t0) > pa.wait_ticks) {
            __hip_atomic_store(pa.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(8);
    }
}

// PERSIST: the epoch is item (pe, task) of a persistent launch -- it counts itself in
// pa->reads[pe] once past its mailbox read, and waits for every item of epoch pe - 1 to
// be past theirs before it overwrites their mailbox buffer.
template <int KB, int FUNC, int ALGO, bool PERSIST = false>
__device__ __forceinline__ void epoch_body(const EpochArgs& a, float* smem, const int task, const TaskDev& td,
                                           const PersistArgs* pa = nullptr, int pe = 0, int pg0 = 0, int pg1 = 0,
                                           int ppub = 0) {
    // the epoch's generations and publish flag: the launch's, or the persistent item's
    const int eg0 = PERSIST ? pg0 : a.g0, eg1 = PERSIST ? pg1 : a.g1, epub = PERSIST ? ppub : a.publish;
    using S = Shape<KB>;
    constexpr int RS = S::RS, HS = S::HS, NT = S::NT;
    const int tid = threadIdx.x;
    Geo g;
    g.D = td.dim;
    g.Dh = (g.D + 1) >> 1;
    g.pop = td.pop;
    g.popP = (g.pop + 31) & ~31;
    g.ntiles = g.popP / 32;
    g.m = g.pop / (a.p.n_islands > 0 ? a.p.n_islands : 1);
    g.modD.init((uint32_t)g.D);
    g.modM.init((uint32_t)g.m);
    Lds L;
    L.X = smem;
    L.Q = L.X + g.popP * RS;
    L.o = L.Q + 32 * NT * RS;
    L.tab = L.o + RS;
    L.fit = L.tab + 32 * NT;
    L.ftmp = L.fit + g.popP;
    L.ord = reinterpret_cast<int*>(L.ftmp + g.popP);
    L.aux = L.ord + g.popP;
    const Rng rng{a.p.seed_lo, a.p.seed_hi, (uint32_t)task};
    const int lane = tid & 63, col = lane & 31, h = lane >> 5, wave = tid >> 6;

    const int nz = L.fit - smem;  // X, Q, o, tab
    for (int e = tid; e < nz; e += kThreads) smem[e] = 0.0f;
    __syncthreads();
    stage_rows<RS, HS>(L.Q, a.Q + td.q_off, g.D, g.D, tid);
    for (int d = tid; d < g.D; d += kThreads) {
        L.o[(d & 1) * HS + (d >> 1)] = a.shift[td.v_off + d];
        L.tab[d] = a.tab[td.v_off + d];
    }
    if (a.init) {
        const float lo = td.lo, rng_ = td.hi - td.lo;
        for (int e = tid; e < g.pop * g.D; e += kThreads) {
            const int i = e / g.D, d = e - i * g.D;
            const float u = uniform01(rng.word(0u, P_INIT, (uint32_t)i, 0u, (uint32_t)d));
            const float t1 = u * rng_;
            L.X[i * RS + (d & 1) * HS + (d >> 1)] = t1 + lo;
        }
        __syncthreads();
        for (int it = wave; it < g.ntiles; it += kThreads / 64) {
            const int i = it * 32 + col;
            float x[HS];
            load_half_row<KB>(x, L.X + i * RS + h * HS);
            const float f = evaluate_tile<KB, FUNC>(x, g, L, col, h);
            if (i < g.pop && h == 0) L.fit[i] = f;
        }
    } else {
        stage_rows<RS, HS>(L.X, a.X + td.x_off, g.pop, g.D, tid);
        for (int i = tid; i < g.pop; i += kThreads) L.fit[i] = a.fit[td.fit_off + i];
    }
    __syncthreads();

    const int TI = a.p.transfer_interval;
    if (transfer_now(a, td, eg0)) transfer_in<KB, FUNC>(a, td, g, L, rng, tid, eg0);
    if constexpr (PERSIST) {
        __syncthreads();
        if (tid == 0) __hip_atomic_fetch_add(pa->reads + pe, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }

    const bool islands = a.p.n_islands > 1;
    for (int gen = eg0 - td.gen_base + 1; gen <= eg1 - td.gen_base; ++gen) {
        if (islands && a.p.exchange_interval > 0 && gen % a.p.exchange_interval == 0)
            island_exchange<KB>(a, g, L, a.scratch + td.scr_off, rng, gen, tid);
        generation_step<KB, FUNC, ALGO>(a, td, g, L, rng, gen, tid);
        if (islands && a.p.migrate_interval > 0 && a.p.migrate_count > 0 && gen % a.p.migrate_interval == 0)
            island_migration<KB>(a, g, L, tid);
    }

    float* Xg = a.X + td.x_off;
    for (int e = tid; e < g.pop * g.D; e += kThreads) {
        const int i = e / g.D, d = e - i * g.D;
        Xg[e] = L.X[i * RS + (d & 1) * HS + (d >> 1)];
    }
    for (int i = tid; i < g.pop; i += kThreads) a.fit[td.fit_off + i] = L.fit[i];
    if (epub) {
        __syncthreads();
        rank_islands(L, g.pop, g.pop, tid);
        if constexpr (PERSIST) {
            __shared__ int s_pub;
            if (tid == 0) s_pub = (pe == 0 || persist_wait(pa->reads + pe - 1, a.n_tasks, *pa)) ? 1 : 0;
            __syncthreads();
            if (!s_pub) return;  // the error flag is set: the launch drains
        }
        __syncthreads();
        const int k = a.p.transfer_count;
        float* mb = a.mailbox + (int64_t)((eg1 / TI) & 1) * a.mailbox_stride + td.mb_off;
        for (int e = tid; e < k * g.D; e += kThreads) {
            const int r = e / g.D, d = e - r * g.D;
            mb[e] = L.X[L.ord[r] * RS + (d & 1) * HS + (d >> 1)];
        }
    }
}

// One workgroup per task; `order` lists the tasks longest-first so the
// expensive objectives are dispatched before the cheap ones fill the gaps.
// Register budget: with KB <= 8 (dims <= 64) three 4-wave workgroups fit in LDS per
// CU, so the allocator is asked for 3 waves per SIMD (<= 168 VGPR+AGPR); larger
// rows are LDS-bound to fewer workgroups and keep the default budget.
#if !defined(EMTF_WAVES_PER_EU) && defined(EMTF_KB) && EMTF_KB <= 8
#define EMTF_WAVES_PER_EU 3
#endif
#ifdef EMTF_WAVES_PER_EU
#define EMTF_OCCUPANCY __attribute__((amdgpu_waves_per_eu(EMTF_WAVES_PER_EU, EMTF_WAVES_PER_EU)))
#else
#define EMTF_OCCUPANCY
#endif

template <int KB, int ALGO>
__global__ __launch_bounds__(kThreads) EMTF_OCCUPANCY void island_epoch_kernel(EpochArgs a, const int* __restrict__ order) {
    extern __shared__ float smem[];
    rebase_epoch(a);
    const int task = order[blockIdx.x];
    const TaskDev td = a.tasks[task];
    if (!runs_epoch(a, td)) return;
    switch (td.func) {
    case F_SPHERE: epoch_body<KB, F_SPHERE, ALGO>(a, smem, task, td); break;
    case F_ROSENBROCK: epoch_body<KB, F_ROSENBROCK, ALGO>(a, smem, task, td); break;
    case F_ACKLEY: epoch_body<KB, F_ACKLEY, ALGO>(a, smem, task, td); break;
    case F_RASTRIGIN: epoch_body<KB, F_RASTRIGIN, ALGO>(a, smem, task, td); break;
    case F_GRIEWANK: epoch_body<KB, F_GRIEWANK, ALGO>(a, smem, task, td); break;
    case F_WEIERSTRASS: epoch_body<KB, F_WEIERSTRASS, ALGO>(a, smem, task, td); break;
    case F_SCHWEFEL: epoch_body<KB, F_SCHWEFEL, ALGO>(a, smem, task, td); break;
    case F_ELLIPTIC: epoch_body<KB, F_ELLIPTIC, ALGO>(a, smem, task, td); break;
    case F_REF_ACKLEY: epoch_body<KB, F_REF_ACKLEY, ALGO>(a, smem, task, td); break;
    case F_REF_RASTRIGIN: epoch_body<KB, F_REF_RASTRIGIN, ALGO>(a, smem, task, td); break;
    case F_REF_GRIEWANK: case F_REF_GRIEWANK_GA: epoch_body<KB, F_REF_GRIEWANK, ALGO>(a, smem, task, td); break;
    case F_REF_WEIERSTRASS: epoch_body<KB, F_REF_WEIERSTRASS, ALGO>(a, smem, task, td); break;
    case F_REF_SCHWEFEL: epoch_body<KB, F_SCHWEFEL, ALGO>(a, smem, task, td); break;
    default: break;
    }
}


// The persistent schedule: generations pa.G0+1 .. pa.G1 of every task in one launch of
// n_epochs x n_tasks workgroups.  Each workgroup takes the next item (epoch e, task
// order[i]), e-major, from pa.next in the order the workgroups start (not by blockIdx), so
// every item it may wait on was taken earlier by a workgroup that is running; it waits for
// its task's epoch e - 1 and, at a transfer boundary, for its donor's (the one transfer_in
// draws), runs the epoch (PERSIST hooks: the mailbox read count and the wait before
// publishing) and releases pa.stamp[task] = e + 1.  No wait targets a later item, so the
// launch finishes whatever the number of resident workgroups; every wait is bounded.
template <int KB, int ALGO>
__global__ __launch_bounds__(kThreads) EMTF_OCCUPANCY void island_persistent_kernel(EpochArgs a, PersistArgs pa,
                                                                                   const int* __restrict__ order) {
    extern __shared__ float smem[];
    __shared__ int s_item, s_ok;
    const int tid = threadIdx.x;
    const int TI = a.p.transfer_interval;
    const bool xfer = TI > 0 && a.n_tasks > 1 && a.p.transfer_count > 0;
    const int n_items = pa.n_epochs * a.n_tasks;
    {
        if (tid == 0) s_item = __hip_atomic_fetch_add(pa.next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int item = s_item;
        if (item >= n_items) return;
        const int e = item / a.n_tasks;
        const int task = order[item - e * a.n_tasks];
        int g0, g1, pub;
        if (xfer) {
            const int b0 = pa.G0 / TI;  // epochs end at the transfer boundaries
            g0 = e == 0 ? pa.G0 : (b0 + e) * TI;
            g1 = min((b0 + e + 1) * TI, pa.G1);
            pub = g1 % TI == 0 ? 1 : 0;
        } else {
            g0 = pa.G0;
            g1 = pa.G1;
            pub = 0;
        }
        const TaskDev td = a.tasks[task];
        if (tid == 0) {
            bool ok = persist_wait(pa.stamp + task, e, pa);
            if (ok && transfer_now(a, td, g0)) {
                const Rng rng{a.p.seed_lo, a.p.seed_hi, (uint32_t)task};
                const uint32_t wd = rng.word((uint32_t)(g0 - td.gen_base), P_MIG, 0u, 0u, 0u);
                ok = persist_wait(pa.stamp + pick_donor(a, task, wd), e, pa);
            }
            s_ok = ok ? 1 : 0;
        }
        __syncthreads();
        if (!s_ok) return;  // the error flag is set: the launch drains
        // (the waits' acquire loads and the stamp's release store, with the workgroup
        // barriers around them, order the rows, fitness and mailbox across workgroups)
        switch (td.func) {
        case F_SPHERE: epoch_body<KB, F_SPHERE, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_ROSENBROCK: epoch_body<KB, F_ROSENBROCK, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_ACKLEY: epoch_body<KB, F_ACKLEY, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_RASTRIGIN: epoch_body<KB, F_RASTRIGIN, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_GRIEWANK: epoch_body<KB, F_GRIEWANK, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_WEIERSTRASS: epoch_body<KB, F_WEIERSTRASS, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_SCHWEFEL: epoch_body<KB, F_SCHWEFEL, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_ELLIPTIC: epoch_body<KB, F_ELLIPTIC, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_REF_ACKLEY: epoch_body<KB, F_REF_ACKLEY, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_REF_RASTRIGIN: epoch_body<KB, F_REF_RASTRIGIN, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_REF_GRIEWANK: case F_REF_GRIEWANK_GA:
            epoch_body<KB, F_REF_GRIEWANK, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub);
            break;
        case F_REF_WEIERSTRASS: epoch_body<KB, F_REF_WEIERSTRASS, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        case F_REF_SCHWEFEL: epoch_body<KB, F_SCHWEFEL, ALGO, true>(a, smem, task, td, &pa, e, g0, g1, pub); break;
        default: break;
        }
        __syncthreads();
        if (tid == 0 && __hip_atomic_load(pa.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
            __hip_atomic_store(pa.stamp + task, e + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int KB, int FUNC, bool ADJ>
__global__ __launch_bounds__(2 * ObjShape<KB>::PB) void objective_kernel(ObjArgs a) {
    using S = Shape<KB>;
    constexpr int RS = S::RS, HS = S::HS, NT = S::NT, PB = ObjShape<KB>::PB, ZS = ObjShape<KB>::ZS;
    extern __shared__ float smem[];
    float* Q = smem;
    float* o = Q + 32 * NT * RS;
    float* tab = o + RS;
    float* Zs = tab + 32 * NT;
    float* Pre = Zs + PB * ZS;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int D = a.dim;
    for (int e = tid; e < 32 * NT * RS + RS + 32 * NT; e += nthr) smem[e] = 0.0f;
    __syncthreads();
    for (int e = tid; e < D * D; e += nthr) {
        const int j = e / D, d = e - j * D;
        Q[j * RS + (d & 1) * HS + (d >> 1)] = a.has_rot ? a.rot[e] : (j == d ? 1.0f : 0.0f);
    }
    for (int d = tid; d < D; d += nthr) {
        o[(d & 1) * HS + (d >> 1)] = a.shift ? a.shift[d] : 0.0f;
        tab[d] = a.tab[d];
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63, col = lane & 31, h = lane >> 5;
    const int pt = blockIdx.x * PB + wave * 32 + col;
    const bool valid = pt < a.n;
    const float* xr = a.x + (int64_t)(valid ? pt : 0) * D;
    float x[HS];
#pragma unroll
    for (int s = 0; s < HS; ++s) {
        const int d = 2 * s + h;
        x[s] = (d < D) ? xr[d] : 0.0f;
    }
    Geo g{};
    g.D = D;
    g.Dh = (D + 1) >> 1;
    Lds L;
    L.Q = Q;
    L.o = o;
    L.tab = tab;
    if constexpr (!ADJ) {
        const float f = evaluate_tile<KB, FUNC>(x, g, L, col, h);
        if (valid && h == 0) a.f[pt] = f;
    } else {
        // z = Q (x - o) as in the forward pass, staged row-wise in LDS
        f16v acc[NT];
#pragma unroll
        for (int jt = 0; jt < NT; ++jt)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[jt][r] = 0.0f;
#pragma unroll
        for (int sb = 0; sb < KB; ++sb) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int s = 4 * sb + e;
                const float w = x[s] - o[h * HS + s];
#pragma unroll
                for (int jt = 0; jt < NT; ++jt)
                    acc[jt] = __builtin_amdgcn_mfma_f32_32x32x2f32(Q[(32 * jt + col) * RS + h * HS + s], w, acc[jt], 0, 0, 0);
            }
        }
        float* zrow = Zs + (wave * 32 + col) * ZS;
#pragma unroll
        for (int jt = 0; jt < NT; ++jt)
#pragma unroll
            for (int r = 0; r < 16; ++r) zrow[32 * jt + 8 * (r >> 2) + 4 * h + (r & 3)] = acc[jt][r];
        __syncthreads();
        float dz[HS];
        dfdz<KB, FUNC>(dz, zrow, Pre + (wave * 32 + col) * ZS, D, tab, h);
        const float up = valid ? a.gf[pt] : 0.0f;
#pragma unroll
        for (int s = 0; s < HS; ++s) dz[s] = dz[s] * up;
        // gx = Q^T dz: A[row d][k j] = Q[j][d]
        f16v gacc[NT];
#pragma unroll
        for (int dt = 0; dt < NT; ++dt)
#pragma unroll
            for (int r = 0; r < 16; ++r) gacc[dt][r] = 0.0f;
        const int dcol = col;  // lane's A row within the tile
#pragma unroll
        for (int s = 0; s < HS; ++s) {
            const int j = 2 * s + h;
#pragma unroll
            for (int dt = 0; dt < NT; ++dt) {
                const int d = 32 * dt + dcol;
                const float qa = (j < 32 * NT) ? Q[j * RS + (d & 1) * HS + (d >> 1)] : 0.0f;
                gacc[dt] = __builtin_amdgcn_mfma_f32_32x32x2f32(qa, dz[s], gacc[dt], 0, 0, 0);
            }
        }
        if (valid) {
            float* gr = a.gx + (int64_t)pt * D;
#pragma unroll
            for (int dt = 0; dt < NT; ++dt)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int d = 32 * dt + 8 * (r >> 2) + 4 * h + (r & 3);
                    if (d < D) gr[d] = gacc[dt][r];
                }
        }
    }
}

#if defined(EMTF_WIDE_NT)
// ---------------------------------------------------------------------------
// Wide tasks: dimensions up to 32 NT whose rows (or whose population) do not fit
// LDS.  The population rows X, the trial rows T and the task's padded Q image
// live in global memory (a task's working set stays L2-resident while its
// workgroup runs); LDS keeps o, the per-dimension table and the fitness and
// rank arrays.  Trial genes are formed block by block and fed straight into the
// MFMA k-loop as the B operand, and stored to T for the selection copy, so no
// lane holds a whole row in registers.  Rows use the LDS path's [even | odd]
// layout with HS = 16 NT; the gene order of every fma chain is the LDS path's,
// with the padded k-steps past the dimension skipped (they would add +0).
// ---------------------------------------------------------------------------
template <int NT>
using WShape = Shape<4 * NT>;

// Q operands of gene block sb: rows 32 jt + col of the lane's half
template <int NT>
__device__ __forceinline__ void load_q(float4 (&q)[NT], const float* qrow, int sb) {
    using S = WShape<NT>;
#pragma unroll
    for (int jt = 0; jt < NT; ++jt) q[jt] = lds4(qrow + 32 * jt * S::RS + 4 * sb);
}

// acc[jt] += Q rows (32 jt + col) x (4 genes of one block of the lane's half)
template <int NT>
__device__ __forceinline__ void mfma_q(f16v (&acc)[NT], const float4 (&q)[NT], const float4& w4) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int jt = 0; jt < NT; ++jt)
            acc[jt] = __builtin_amdgcn_mfma_f32_32x32x2f32(f4(q[jt], e), f4(w4, e), acc[jt], 0, 0, 0);
}

template <int NT>
__device__ __forceinline__ void mfma_block(f16v (&acc)[NT], const float* qrow, const float4& w4, int sb) {
    float4 q[NT];
    load_q<NT>(q, qrow, sb);
    mfma_q<NT>(acc, q, w4);
}

__device__ __forceinline__ float4 sub4(const float4& x, const float4& o) {
    return make_float4(x.x - o.x, x.y - o.y, x.z - o.z, x.w - o.w);
}

template <int NT>
__device__ __forceinline__ void zero_acc(f16v (&acc)[NT]) {
#pragma unroll
    for (int jt = 0; jt < NT; ++jt)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[jt][r] = 0.0f;
}

// f of the lane's row (its parity half at row_h) read from memory
template <int NT, int FUNC>
__device__ __forceinline__ float wide_eval_row(const float* row_h, const Geo& g, const Lds& L, int col, int h, int nsb) {
    using S = WShape<NT>;
    f16v acc[NT];
    zero_acc<NT>(acc);
    const float* qrow = L.Q + col * S::RS + h * S::HS;
    const float* orow = L.o + h * S::HS;
    for (int sb = 0; sb < nsb; ++sb) mfma_block<NT>(acc, qrow, sub4(lds4(row_h + 4 * sb), lds4(orow + 4 * sb)), sb);
    return objective<NT, FUNC, true>(acc, g.D, L.tab, h);
}

template <int NT, int FUNC, int ALGO, int WT, bool QL>
__device__ __forceinline__ void wide_generation(const EpochArgs& a, const TaskDev& td, const Geo& g, const Lds& L,
                                                float* T, const Rng& rng, int gen, int tid, int nsb) {
    using S = WShape<NT>;
    constexpr int RS = S::RS, HS = S::HS;
    const int wave = tid >> 6, lane = tid & 63, col = lane & 31;
    for (int it = wave; it < g.ntiles; it += WT / 64) {
        // the lane's half index is made opaque per tile (masks derived from it are not
        // hoisted out of the generation loop into spilled scalar registers)
        int h = lane >> 5;
        asm volatile("" : "+v"(h));
        const float* qrow = L.Q + col * RS + h * HS;
        const float* orow = L.o + h * HS;
        const int i = it * 32 + col;
        const bool valid = i < g.pop;
        const int ie = valid ? i : 0;
        const int il = ie % g.m, base = ie - il;
        float* trow = T + (int64_t)i * RS + h * HS;
        f16v acc[NT];
        zero_acc<NT>(acc);
        if constexpr (ALGO == A_GA_SBX_PM) {
            const U4 mb = rng.block((uint32_t)gen, P_GA_MATE, (uint32_t)(ie & ~1), 0u, 0u);
            const int ra = (int)g.modM.mod(pick(mb, (uint32_t)(il & 1)));
            const int rb = (int)g.modM.mod(pick(mb, (uint32_t)((il & 1) ^ 1)));
            const float* xa = L.X + (int64_t)(base + ra) * RS + h * HS;
            const float* xb = L.X + (int64_t)(base + rb) * RS + h * HS;
            const GaConst c0 = ga_const(a, td);
            const bool odd = (il & 1) != 0;
            float4 na = lds4(xa), nb = lds4(xb);
            float4 nq[NT];
            if constexpr (!QL) load_q<NT>(nq, qrow, 0);
            for (int sb = 0; sb < nsb; ++sb) {
                const float4 va = na, vb = nb;
                float4 q[NT];
                if constexpr (QL) {
                    load_q<NT>(q, qrow, sb);
                } else {
#pragma unroll
                    for (int jt = 0; jt < NT; ++jt) q[jt] = nq[jt];
                }
                if (sb + 1 < nsb) {
                    na = lds4(xa + 4 * (sb + 1));
                    nb = lds4(xb + 4 * (sb + 1));
                    if constexpr (!QL) load_q<NT>(nq, qrow, sb + 1);
                }
                float t[4];
                ga_block(t, va, vb, rng, gen, ie, sb, h, odd, g.D, c0);
                const float4 t4 = make_float4(t[0], t[1], t[2], t[3]);
                *reinterpret_cast<float4*>(trow + 4 * sb) = t4;
                mfma_q<NT>(acc, q, sub4(t4, lds4(orow + 4 * sb)));
            }
        } else {
            const bool rand1 = (a.p.algo == A_DE_RAND_1_BIN);
            const DeMates mt = de_mates(a, g, rng, gen, ie);
            const float* xb = L.X + (int64_t)(base + (rand1 ? mt.r1 : il)) * RS + h * HS;
            const float* x2 = L.X + (int64_t)(base + mt.r2) * RS + h * HS;
            const float* x3 = L.X + (int64_t)(base + mt.r3) * RS + h * HS;
            const float* xi = L.X + (int64_t)ie * RS + h * HS;
            // global-memory operands of block sb + 1 are loaded while block sb computes
            U4 cb{0u, 0u, 0u, 0u};
            float4 nb = lds4(xb), n2 = lds4(x2), n3 = lds4(x3), ni = lds4(xi);
            float4 nq[NT];
            if constexpr (!QL) load_q<NT>(nq, qrow, 0);
            for (int sb = 0; sb < nsb; ++sb) {
                const float4 vb = nb, v2 = n2, v3 = n3, vi = ni;
                float4 q[NT];
                if constexpr (QL) {
                    load_q<NT>(q, qrow, sb);
                } else {
#pragma unroll
                    for (int jt = 0; jt < NT; ++jt) q[jt] = nq[jt];
                }
                if (sb + 1 < nsb) {
                    const int o = 4 * (sb + 1);
                    nb = lds4(xb + o);
                    n2 = lds4(x2 + o);
                    n3 = lds4(x3 + o);
                    ni = lds4(xi + o);
                    if constexpr (!QL) load_q<NT>(nq, qrow, sb + 1);
                }
                if ((sb & 1) == 0) cb = rng.block((uint32_t)gen, P_DE_CR, (uint32_t)ie, (uint32_t)h, (uint32_t)(sb >> 1));
                float t[4];
                de_block(t, vb, v2, v3, vi, cb, sb, h, mt.jrand, a.p.F, td.lo, td.hi, a.p.cr_threshold);
                const float4 t4 = make_float4(t[0], t[1], t[2], t[3]);
                *reinterpret_cast<float4*>(trow + 4 * sb) = t4;
                mfma_q<NT>(acc, q, sub4(t4, lds4(orow + 4 * sb)));
            }
        }
        const float f = objective<NT, FUNC, true>(acc, g.D, L.tab, h);
        const bool accept = valid && (L.fit[i] >= f);
        if (valid && h == 0) {
            L.aux[i] = accept ? 1 : 0;
            L.ftmp[i] = f;
        }
    }
    __syncthreads();
    for (int e = tid; e < g.pop * 2 * nsb; e += WT) {
        const int row = e / (2 * nsb), blk = e - row * (2 * nsb);
        if (L.aux[row]) {
            const int c = (blk < nsb) ? 4 * blk : HS + 4 * (blk - nsb);
            *reinterpret_cast<float4*>(L.X + (int64_t)row * RS + c) = lds4(T + (int64_t)row * RS + c);
        }
    }
    for (int i = tid; i < g.pop; i += WT)
        if (L.aux[i]) L.fit[i] = L.ftmp[i];
    __syncthreads();
}

// Inter-task transfer for wide tasks: candidates streamed into the k-loop and
// staged in T (row j for candidate j), accepted ones copied to their target row.
template <int NT, int FUNC, int WT>
__device__ __forceinline__ void wide_transfer_in(const EpochArgs& a, const TaskDev& td, const Geo& g, const Lds& L,
                                                 float* T, const Rng& rng, int tid, int nsb) {
    using S = WShape<NT>;
    constexpr int RS = S::RS, HS = S::HS;
    const int k = a.p.transfer_count, TI = a.p.transfer_interval, g0 = a.g0 - td.gen_base;
    const int wave = tid >> 6, lane = tid & 63, col = lane & 31, h = lane >> 5;
    rank_islands<WT>(L, g.pop, g.pop, tid);
    __syncthreads();
    const uint32_t wd = rng.word((uint32_t)g0, P_MIG, 0u, 0u, 0u);
    const int dn = pick_donor(a, (int)rng.task, wd);
    const TaskDev dd = a.tasks[dn];
    const float* mb = a.mailbox + (int64_t)((a.g0 / TI) & 1) * a.mailbox_stride + dd.mb_off;
    const bool same_box = (dd.lo == td.lo && dd.hi == td.hi);
    const float rng_s = td.hi - td.lo, rng_d = dd.hi - dd.lo;
    const float* qrow = L.Q + col * RS + h * HS;
    const float* orow = L.o + h * HS;
    const int ntc = (k + 31) / 32;
    for (int ct = wave; ct < ntc; ct += WT / 64) {
        const int j = ct * 32 + col;
        const bool valid = j < k;
        const int je = valid ? j : 0;
        const float* src = L.X + (int64_t)L.ord[je] * RS + h * HS;
        float* crow = T + (int64_t)j * RS + h * HS;
        f16v acc[NT];
        zero_acc<NT>(acc);
        for (int sb = 0; sb < nsb; ++sb) {
            const float4 v4 = lds4(src + 4 * sb);
            float c[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int s = 4 * sb + e, d = 2 * s + h;
                if (d < g.D && d < dd.dim) {
                    bool take = true;
                    if (a.p.transfer_cr > 0.0f) {
                        const uint32_t wu = rng.word((uint32_t)g0, P_XFER, (uint32_t)je, (uint32_t)h, (uint32_t)s);
                        take = uniform01(wu) > a.p.transfer_cr;
                    }
                    if (take) {
                        const float v = mb[(int64_t)je * dd.dim + d];
                        if (same_box) c[e] = v;
                        else {
                            const float t1 = v - dd.lo;
                            const float t2 = t1 / rng_d;
                            const float t3 = t2 * rng_s;
                            c[e] = td.lo + t3;
                        }
                    }
                }
            }
            const float4 c4 = make_float4(c[0], c[1], c[2], c[3]);
            *reinterpret_cast<float4*>(crow + 4 * sb) = c4;
            mfma_block<NT>(acc, qrow, sub4(c4, lds4(orow + 4 * sb)), sb);
        }
        const float f = objective<NT, FUNC, true>(acc, g.D, L.tab, h);
        if (valid) {
            const int tgt = L.ord[g.pop - k + j];
            if (f < L.fit[tgt]) {
                float* dst = L.X + (int64_t)tgt * RS + h * HS;
                for (int sb = 0; sb < nsb; ++sb) *reinterpret_cast<float4*>(dst + 4 * sb) = lds4(crow + 4 * sb);
                if (h == 0) L.fit[tgt] = f;
            }
        }
    }
    __syncthreads();
}

// QL: the task's Q image is staged in LDS (the workgroup's first 32 NT x RS floats) and
// the MFMA A operands come from there instead of L2.
template <int NT, int FUNC, int ALGO, int WT, bool QL>
__device__ __forceinline__ void wide_body(const EpochArgs& a, float* smem, const int task, const TaskDev& td) {
    using S = WShape<NT>;
    constexpr int RS = S::RS, HS = S::HS;
    const int tid = threadIdx.x;
    Geo g;
    g.D = td.dim;
    g.Dh = (g.D + 1) >> 1;
    g.pop = td.pop;
    g.popP = (g.pop + 31) & ~31;
    g.ntiles = g.popP / 32;
    g.m = g.pop / (a.p.n_islands > 0 ? a.p.n_islands : 1);
    g.modD.init((uint32_t)g.D);
    g.modM.init((uint32_t)g.m);
    const int nsb = (g.Dh + 3) >> 2;  // gene blocks per half that hold genes
    Lds L;
    L.X = a.scratch + td.scr_off;
    float* T = L.X + (int64_t)g.popP * RS;
    float* base = smem;
    if constexpr (QL) {
        L.Q = smem;
        base = smem + 32 * NT * RS;
        // eight 16-byte loads in flight per thread
        const float4* qg = reinterpret_cast<const float4*>(a.Qw + td.qw_off);
        constexpr int n4 = 32 * NT * RS / 4;
        for (int e0 = tid; e0 < n4; e0 += 8 * WT) {
            float4 v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int e = e0 + k * WT;
                v[k] = qg[e < n4 ? e : n4 - 1];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int e = e0 + k * WT;
                if (e < n4) reinterpret_cast<float4*>(smem)[e] = v[k];
            }
        }
    } else {
        L.Q = const_cast<float*>(a.Qw + td.qw_off);
    }
    L.o = base;
    L.tab = L.o + RS;
    L.fit = L.tab + 32 * NT;
    L.ftmp = L.fit + g.popP;
    L.ord = reinterpret_cast<int*>(L.ftmp + g.popP);
    L.aux = L.ord + g.popP;
    const Rng rng{a.p.seed_lo, a.p.seed_hi, (uint32_t)task};
    const int lane = tid & 63, col = lane & 31, h = lane >> 5, wave = tid >> 6;

    for (int e = tid; e < RS + 32 * NT; e += WT) base[e] = 0.0f;
    __syncthreads();
    for (int d = tid; d < g.D; d += WT) {
        L.o[(d & 1) * HS + (d >> 1)] = a.shift[td.v_off + d];
        L.tab[d] = a.tab[td.v_off + d];
    }
    // rows in the padded layout, zero past the dimension
    const float* Xg = a.X + td.x_off;
    const float lo = td.lo, rng_ = td.hi - td.lo;
    const int nx = g.popP * RS;
    if (a.init) {
        for (int e = tid; e < nx; e += WT) {
            const int i = e / RS, c = e - i * RS;
            const int d = (c < HS) ? 2 * c : 2 * (c - HS) + 1;
            float v = 0.0f;
            if (i < g.pop && c < 2 * HS && d < g.D) {
                const float u = uniform01(rng.word(0u, P_INIT, (uint32_t)i, 0u, (uint32_t)d));
                const float t1 = u * rng_;
                v = t1 + lo;
            }
            L.X[e] = v;
        }
    } else {
        for (int e0 = tid; e0 < nx; e0 += 8 * WT) {  // eight loads in flight per thread
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int e = e0 + k * WT;
                const int i = e / RS, c = e - i * RS;
                const int d = (c < HS) ? 2 * c : 2 * (c - HS) + 1;
                const bool in = e < nx && i < g.pop && c < 2 * HS && d < g.D;
                const float xv = Xg[in ? (int64_t)i * g.D + d : 0];
                v[k] = in ? xv : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int e = e0 + k * WT;
                if (e < nx) L.X[e] = v[k];
            }
        }
    }
    if (!a.init)
        for (int i = tid; i < g.pop; i += WT) L.fit[i] = a.fit[td.fit_off + i];
    __syncthreads();
    if (a.init) {
        for (int it = wave; it < g.ntiles; it += WT / 64) {
            const int i = it * 32 + col;
            const float f = wide_eval_row<NT, FUNC>(L.X + (int64_t)i * RS + h * HS, g, L, col, h, nsb);
            if (i < g.pop && h == 0) L.fit[i] = f;
        }
        __syncthreads();
    }

    const int TI = a.p.transfer_interval;
    if (transfer_now(a, td)) wide_transfer_in<NT, FUNC, WT>(a, td, g, L, T, rng, tid, nsb);

    const bool islands = a.p.n_islands > 1;
    for (int gen = a.g0 - td.gen_base + 1; gen <= a.g1 - td.gen_base; ++gen) {
        if (islands && a.p.exchange_interval > 0 && gen % a.p.exchange_interval == 0)
            island_exchange<4 * NT, WT>(a, g, L, T, rng, gen, tid);
        wide_generation<NT, FUNC, ALGO, WT, QL>(a, td, g, L, T, rng, gen, tid, nsb);
        if (islands && a.p.migrate_interval > 0 && a.p.migrate_count > 0 && gen % a.p.migrate_interval == 0)
            island_migration<4 * NT, WT>(a, g, L, tid);
    }

    float* Xo = a.X + td.x_off;
    for (int e = tid; e < g.pop * g.D; e += WT) {
        const int i = e / g.D, d = e - i * g.D;
        Xo[e] = L.X[(int64_t)i * RS + (d & 1) * HS + (d >> 1)];
    }
    for (int i = tid; i < g.pop; i += WT) a.fit[td.fit_off + i] = L.fit[i];
    if (a.publish) {
        __syncthreads();
        rank_islands<WT>(L, g.pop, g.pop, tid);
        __syncthreads();
        const int k = a.p.transfer_count;
        float* mb = a.mailbox + (int64_t)((a.g1 / TI) & 1) * a.mailbox_stride + td.mb_off;
        for (int e = tid; e < k * g.D; e += WT) {
            const int r = e / g.D, d = e - r * g.D;
            mb[e] = L.X[(int64_t)L.ord[r] * RS + (d & 1) * HS + (d >> 1)];
        }
    }
}

// Two waves per SIMD: a wave's MFMA k-loop then overlaps another wave's trial
// formation and objective epilogue (one wave per SIMD leaves the MFMA pipe idle
// through both).  Workgroups of 256 threads (two per CU) or, for populations of
// more than four 32-row tiles, 512 threads (one per CU, half the tiles per wave).
// WPE: the waves per SIMD the register allocation targets (1: no target)
constexpr int kWideWpe = EMTF_WIDE_NT <= 6 ? 2 : 1;

template <int NT, int ALGO, int WT, bool QL, int WPE>
__global__ __launch_bounds__(WT) __attribute__((amdgpu_waves_per_eu(WPE))) void wide_epoch_kernel(
    EpochArgs a, const int* __restrict__ order) {
    extern __shared__ float smem[];
    rebase_epoch(a);
    const int task = order[blockIdx.x];
    const TaskDev td = a.tasks[task];
    if (!runs_epoch(a, td)) return;
    switch (td.func) {
    case F_SPHERE: wide_body<NT, F_SPHERE, ALGO, WT, QL>(a, smem, task, td); break;
    case F_ROSENBROCK: wide_body<NT, F_ROSENBROCK, ALGO, WT, QL>(a, smem, task, td); break;
    case F_ACKLEY: wide_body<NT, F_ACKLEY, ALGO, WT, QL>(a, smem, task, td); break;
    case F_RASTRIGIN: wide_body<NT, F_RASTRIGIN, ALGO, WT, QL>(a, smem, task, td); break;
    case F_GRIEWANK: wide_body<NT, F_GRIEWANK, ALGO, WT, QL>(a, smem, task, td); break;
    case F_WEIERSTRASS: wide_body<NT, F_WEIERSTRASS, ALGO, WT, QL>(a, smem, task, td); break;
    case F_SCHWEFEL: wide_body<NT, F_SCHWEFEL, ALGO, WT, QL>(a, smem, task, td); break;
    case F_ELLIPTIC: wide_body<NT, F_ELLIPTIC, ALGO, WT, QL>(a, smem, task, td); break;
    case F_REF_ACKLEY: wide_body<NT, F_REF_ACKLEY, ALGO, WT, QL>(a, smem, task, td); break;
    case F_REF_RASTRIGIN: wide_body<NT, F_REF_RASTRIGIN, ALGO, WT, QL>(a, smem, task, td); break;
    case F_REF_GRIEWANK: case F_REF_GRIEWANK_GA: wide_body<NT, F_REF_GRIEWANK, ALGO, WT, QL>(a, smem, task, td); break;
    case F_REF_WEIERSTRASS: wide_body<NT, F_REF_WEIERSTRASS, ALGO, WT, QL>(a, smem, task, td); break;
    case F_REF_SCHWEFEL: wide_body<NT, F_SCHWEFEL, ALGO, WT, QL>(a, smem, task, td); break;
    default: break;
    }
}

// emtf_objective_fwd for dimensions past the LDS evaluator's (D <= 32 NT): Q and the
// points are read row-major from global memory, one gene pair per MFMA k-step in
// increasing order with zero padding past D -- the engine evaluator's fma chains and
// reductions, so the same bits.  128 points per workgroup, 32 per wave.
template <int NT, int FUNC>
__global__ __launch_bounds__(256) void objective_wide_kernel(ObjArgs a) {
    __shared__ float o_s[32 * NT], tab_s[32 * NT];
    const int tid = threadIdx.x, D = a.dim;
    for (int d = tid; d < 32 * NT; d += 256) {
        o_s[d] = (d < D && a.shift) ? a.shift[d] : 0.0f;
        tab_s[d] = d < D ? a.tab[d] : 0.0f;
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63, col = lane & 31, h = lane >> 5;
    const int pt = blockIdx.x * 128 + wave * 32 + col;
    const bool valid = pt < a.n;
    const float* xr = a.x + (int64_t)(valid ? pt : 0) * D;
    f16v acc[NT];
    zero_acc<NT>(acc);
    const int Dh = (D + 1) >> 1;
    for (int p = 0; p < Dh; ++p) {
        const int d = 2 * p + h;
        const float w = d < D ? xr[d] - o_s[d] : 0.0f;
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) {
            const int row = 32 * jt + col;
            float q = 0.0f;
            if (row < D && d < D) q = a.has_rot ? a.rot[(int64_t)row * D + d] : (row == d ? 1.0f : 0.0f);
            acc[jt] = __builtin_amdgcn_mfma_f32_32x32x2f32(q, w, acc[jt], 0, 0, 0);
        }
    }
    const float f = objective<NT, FUNC, true>(acc, D, tab_s, h);
    if (valid && h == 0) a.f[pt] = f;
}
#endif  // EMTF_WIDE_NT

}  // namespace emtf

#if defined(EMTF_WIDE_NT)
#if !defined(EMTF_ALGO)
#error "compile the wide kernel with -DEMTF_WIDE_NT=<row tiles> -DEMTF_ALGO=<0|2>"
#endif
#define EMTF_WCAT(a, b, c, d) a##b##_##c##d
#define EMTF_WNAME(a, b, c, d) EMTF_WCAT(a, b, c, d)

// Wide-task LDS: o, table, fitness, scratch fitness, ranks, flags.
extern "C" int EMTF_WNAME(emtf_wide_lds_floats_nt, EMTF_WIDE_NT, a, EMTF_ALGO)(int pop) {
    const int popP = (pop + 31) & ~31;
    return emtf::WShape<EMTF_WIDE_NT>::RS + 32 * EMTF_WIDE_NT + 4 * popP;
}

extern "C" int EMTF_WNAME(emtf_launch_wide_nt, EMTF_WIDE_NT, a, EMTF_ALGO)(const emtf::EpochArgs* args, const int* order,
                                                                          int n, int lds_bytes, hipStream_t stream) {
    hipLaunchKernelGGL((emtf::wide_epoch_kernel<EMTF_WIDE_NT, EMTF_ALGO, 256, false, emtf::kWideWpe>), dim3(n), dim3(256), lds_bytes, stream,
                       *args, order);
    return (int)hipGetLastError();
}

extern "C" int EMTF_WNAME(emtf_launch_wide8_nt, EMTF_WIDE_NT, a, EMTF_ALGO)(const emtf::EpochArgs* args, const int* order,
                                                                           int n, int lds_bytes, hipStream_t stream) {
    hipLaunchKernelGGL((emtf::wide_epoch_kernel<EMTF_WIDE_NT, EMTF_ALGO, 512, false, emtf::kWideWpe>), dim3(n), dim3(512), lds_bytes, stream,
                       *args, order);
    return (int)hipGetLastError();
}

// 512 threads with the Q image in LDS (lds_bytes includes it): built where it fits
#if EMTF_WIDE_NT <= 6
extern "C" int EMTF_WNAME(emtf_launch_wideq_nt, EMTF_WIDE_NT, a, EMTF_ALGO)(const emtf::EpochArgs* args, const int* order,
                                                                           int n, int lds_bytes, hipStream_t stream) {
    static bool lds_ok = false;
    auto k = &emtf::wide_epoch_kernel<EMTF_WIDE_NT, EMTF_ALGO, 512, true, emtf::kWideWpe>;
    if (!lds_ok) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return (int)e;
        lds_ok = true;
    }
    hipLaunchKernelGGL(k, dim3(n), dim3(512), lds_bytes, stream, *args, order);
    return (int)hipGetLastError();
}

extern "C" int EMTF_WNAME(emtf_wideq_lds_floats_nt, EMTF_WIDE_NT, a, EMTF_ALGO)(int pop) {
    const int popP = (pop + 31) & ~31;
    using S = emtf::WShape<EMTF_WIDE_NT>;
    return 32 * EMTF_WIDE_NT * S::RS + S::RS + 32 * EMTF_WIDE_NT + 4 * popP;
}
#endif

#if EMTF_ALGO == 0 && EMTF_WIDE_NT >= 5
// emtf_objective_fwd for 128 < dim <= 32 NT (forward only)
template <int FUNC>
static hipError_t launch_obj_wide(const emtf::ObjArgs& a, hipStream_t stream) {
    hipLaunchKernelGGL((emtf::objective_wide_kernel<EMTF_WIDE_NT, FUNC>), dim3((a.n + 127) / 128), dim3(256), 0, stream, a);
    return hipGetLastError();
}

extern "C" int EMTF_WNAME(emtf_launch_objective_nt, EMTF_WIDE_NT, fw, d)(const emtf::ObjArgs* a, hipStream_t s) {
    switch (a->func) {
    case 0: return (int)launch_obj_wide<0>(*a, s);
    case 1: return (int)launch_obj_wide<1>(*a, s);
    case 2: return (int)launch_obj_wide<2>(*a, s);
    case 3: return (int)launch_obj_wide<3>(*a, s);
    case 4: return (int)launch_obj_wide<4>(*a, s);
    case 5: return (int)launch_obj_wide<5>(*a, s);
    case 6: return (int)launch_obj_wide<6>(*a, s);
    case 7: return (int)launch_obj_wide<7>(*a, s);
    }
    return (int)hipErrorInvalidValue;
}
#endif
#else
#if !defined(EMTF_KB) || !defined(EMTF_ALGO)
#error "compile island.hip with -DEMTF_KB=<gene blocks per half> -DEMTF_ALGO=<0|2>"
#endif

#define EMTF_CAT3(a, b, c, d) a##b##_##c##d
#define EMTF_NAME(a, b, c, d) EMTF_CAT3(a, b, c, d)

// emtf_launch_epoch_kb<KB>_a<ALGO>: one instantiation per translation unit.
extern "C" int EMTF_NAME(emtf_launch_epoch_kb, EMTF_KB, a, EMTF_ALGO)(const emtf::EpochArgs* args, const int* order,
                                                                     int n, int lds_bytes, hipStream_t stream) {
    static bool lds_ok = false;
    auto k = &emtf::island_epoch_kernel<EMTF_KB, EMTF_ALGO>;
    if (!lds_ok) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return (int)e;
        lds_ok = true;
    }
    hipLaunchKernelGGL(k, dim3(n), dim3(emtf::kThreads), lds_bytes, stream, *args, order);
    return (int)hipGetLastError();
}

// emtf_launch_persist_kb<KB>_a<ALGO>: the persistent schedule's kernel over `blocks`
// workgroups (one per item).
// the kernel's dynamic LDS allowance: 160 KB less its own static LDS (its item slot)
static hipError_t persist_lds_setup() {
    static bool done = false;
    if (done) return hipSuccess;
    const void* k = reinterpret_cast<const void*>(&emtf::island_persistent_kernel<EMTF_KB, EMTF_ALGO>);
    hipFuncAttributes fa{};
    hipError_t e = hipFuncGetAttributes(&fa, k);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - (int)fa.sharedSizeBytes);
    if (e == hipSuccess) done = true;
    return e;
}

extern "C" int EMTF_NAME(emtf_launch_persist_kb, EMTF_KB, a, EMTF_ALGO)(const emtf::EpochArgs* args,
                                                                       const emtf::PersistArgs* pa, const int* order,
                                                                       int blocks, int lds_bytes, hipStream_t stream) {
    const hipError_t e = persist_lds_setup();
    if (e != hipSuccess) return (int)e;
    auto k = &emtf::island_persistent_kernel<EMTF_KB, EMTF_ALGO>;
    hipLaunchKernelGGL(k, dim3(blocks), dim3(emtf::kThreads), lds_bytes, stream, *args, *pa, order);
    return (int)hipGetLastError();
}


#if EMTF_KB == 2 && EMTF_ALGO == 0
extern "C" int emtf_lds_floats_kb(int pop, int kb) { return emtf::lds_floats_kb(pop, kb); }
#endif

#if EMTF_ALGO == 0
// Objective kernels live in the DE translation unit of each KB.
template <int FUNC, bool ADJ>
static hipError_t launch_obj(const emtf::ObjArgs& a, hipStream_t stream) {
    constexpr int KB = EMTF_KB;
    auto k = &emtf::objective_kernel<KB, FUNC, ADJ>;
    const int bytes = 4 * emtf::obj_lds_floats<KB>();
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    constexpr int PB = emtf::ObjShape<KB>::PB;
    hipLaunchKernelGGL(k, dim3((a.n + PB - 1) / PB), dim3(2 * PB), bytes, stream, a);
    return hipGetLastError();
}

template <bool ADJ>
static hipError_t launch_obj_any(const emtf::ObjArgs& a, hipStream_t s) {
    switch (a.func) {
    case 0: return launch_obj<0, ADJ>(a, s);
    case 1: return launch_obj<1, ADJ>(a, s);
    case 2: return launch_obj<2, ADJ>(a, s);
    case 3: return launch_obj<3, ADJ>(a, s);
    case 4: return launch_obj<4, ADJ>(a, s);
    case 5: return launch_obj<5, ADJ>(a, s);
    case 6: return launch_obj<6, ADJ>(a, s);
    case 7: return launch_obj<7, ADJ>(a, s);
    }
    return hipErrorInvalidValue;
}

#define EMTF_OBJ_NAME2(a, b) a##b
#define EMTF_OBJ_NAME(a, b) EMTF_OBJ_NAME2(a, b)

// Polish kernel LDS: Q, o, tab, fit, ord, then the staging of 32 / cpe elite slots.
extern "C" int EMTF_OBJ_NAME(emtf_polish_lds_floats_kb, EMTF_KB)(int pop, int waves, int cpe) {
    using S = emtf::Shape<EMTF_KB>;
    (void)waves;
    const int popP = (pop + 31) & ~31;
    return 32 * S::NT * S::RS + S::RS + 32 * S::NT + 2 * popP + (32 / cpe) * emtf::PolishSlot<EMTF_KB>::floats(cpe) +
           emtf::PolishSlot<EMTF_KB>::kScratch;
}

extern "C" int EMTF_OBJ_NAME(emtf_launch_polish_kb, EMTF_KB)(const emtf::EpochArgs* args, const int* order, int n,
                                                             int lds_bytes, hipStream_t stream) {
    static bool lds_ok = false;
    auto k = &emtf::polish_kernel<EMTF_KB>;
    if (!lds_ok) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return (int)e;
        lds_ok = true;
    }
    hipLaunchKernelGGL(k, dim3(n), dim3(emtf::kThreads), lds_bytes, stream, *args, order);
    return (int)hipGetLastError();
}
extern "C" int EMTF_OBJ_NAME(emtf_launch_objective_kb, EMTF_KB)(const emtf::ObjArgs* a, int adjoint, hipStream_t s) {
    return (int)(adjoint ? launch_obj_any<true>(*a, s) : launch_obj_any<false>(*a, s));
}
#endif
#endif  // EMTF_WIDE_NT
