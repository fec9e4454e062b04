/*
 * or_engine.c -- the island-model evolutionary-multitasking loop of the oracle.
 * TEST INFRASTRUCTURE ONLY (see emtf_oracle.h); also timed as the CPU
 * baseline (OpenMP across tasks).
 *
 * Restates, per generation and per task (reference file:line):
 *   init           pop_init            proposed model/SG-EMTF.cu:146-159 (x = U*range + lo)
 *   DE mutant      CrossOver           SG-EMTF_DE.cu:547-593 (distinct r1,r2,r3 by rejection
 *                                      on curand()%n; v = base + F (x_r2 - x_r3); clamp)
 *   DE binomial    Mutation            SG-EMTF_DE.cu:474-496 (jrand = curand()%D; keep the
 *                                      parent gene iff u > CR and d != jrand)
 *   GA mate index  CrossPrep           SG-EMTF.cu:803-813 (r[i] = curand()%n)
 *   GA SBX         CrossOver           SG-EMTF.cu:711-749 (beta from p, negated for odd i,
 *                                      parents r[i], r[i^1]; clamp)
 *   GA polynomial  Mutation            SG-EMTF.cu:548-606 (Deb's two-sided delta; clamp)
 *   selection      Selection           SG-EMTF.cu:752-769, SG-EMTF_DE.cu:679-705 (child
 *                                      replaces parent iff f_parent >= f_child)
 *   island exchange SharedToGlobal + SelectFromBlocks  SG-EMTF.cu:856-896, :1220-1260
 *                                      (Fisher-Yates shuffle with curand()%(k+1), then
 *                                      position k refilled from island (q + r) mod n)
 *   island migration IslandMigration + paritySort  SG-EMTF_DE.cu:1072-1170 (rank j of the
 *                                      source replaces rank m-k+j of the destination iff
 *                                      strictly better)
 *   transfer       knowledgeTransfer   SG-EMTF.cu:1514-1591 (+ partner choice :1696-1699):
 *                                      gene taken from the partner's rank-j elite, mapped
 *                                      box to box, iff u > CR; evaluated on self; replaces
 *                                      self's rank pop-k+j iff strictly better
 * with the draw addressing and arithmetic contract of emtf_oracle.h.
 */
#include "emtf_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Planted mistakes for the mutation tests (tests/test_oracle_pins.py builds this file with
 * -DOR_MUTANT=k into a temporary directory and checks that the property tests reject the
 * result).  The oracle itself is built with OR_MUTANT undefined, i.e. 0: every macro below
 * is then the plain expression on its right. */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif
#define MUT_XCH_INDEX(q) (OR_MUTANT == 1 ? (uint32_t)(q) : 0u)               /* offset per (island, slot) */
#define MUT_MIG_DST(m, k, j) (OR_MUTANT == 2 ? (m) - 1 - (j) : (m) - (k) + (j)) /* destination rank */
#define MUT_MIG_BETTER(a, b) (OR_MUTANT == 3 ? (a) <= (b) : (a) < (b))          /* strictly better */
#define MUT_MIG_SRC(m, j) (OR_MUTANT == 4 ? (m) - 1 - (j) : (j))               /* source rank */
#define MUT_XFER_RD(rd, rs) (OR_MUTANT == 5 ? (rs) : (rd))                    /* box map ranges swapped */
#define MUT_XFER_RS(rd, rs) (OR_MUTANT == 5 ? (rd) : (rs))
#define MUT_XFER_TGT(P, k, j) (OR_MUTANT == 6 ? (j) : (P) - (k) + (j))         /* replaced rank */
#define MUT_DONOR_END(end, g) (OR_MUTANT == 8 ? (end) >= (g) : (end) > (g))     /* done at g still a donor */
#define MUT_DONOR_BASE(base, g) (OR_MUTANT == 9 ? (base) <= (g) : (base) < (g)) /* new at g already a donor */
#define MUT_DONOR_POS(pos, ti) (OR_MUTANT == 10 ? (ti) : (pos))                 /* skip self by id */
#define MUT_LOCAL_GEN(g, base) (OR_MUTANT == 11 ? (g) : (g) - (base))           /* draws by the clock */

typedef struct {
    or_task t;          /* pointers below are owned copies */
    float* rot;         /* Q[j][d] */
    float* rotT;        /* Q[d][j] (for the vectorised GEMV) */
    float* shift;
    float* diag;
    float* tab;
    float* X;           /* pop * D */
    float* fit;         /* pop */
    /* scratch */
    float* trial;       /* pop * D */
    float* tfit;        /* pop */
    float* snapX;       /* pop * D */
    float* snapF;       /* pop */
    int* order;         /* pop */
    float* elites;      /* transfer_count * D (snapshot for transfer) */
    /* life on the engine clock: runs generations base + 1 .. end, its own generation
     * counter (the one its draws are addressed by) is the clock minus base */
    int base, end;
} task_state;

#define OR_NEVER 0x7FFFFFFF

struct or_engine {
    int T;
    or_params p;
    int gen;
    task_state* ts;
    int* donors;        /* scratch: the donor list of a transfer event */
};

static inline float uniform01(uint32_t w) { return (float)((w >> 8) + 1u) * 5.9604644775390625e-08f; }
/* 16-bit uniform in (0,1]: half 0 = high 16 bits, half 1 = low 16 bits */
static inline float uniform16(uint32_t w, int half) {
    uint32_t v = half ? (w & 0xFFFFu) : (w >> 16);
    return (float)(v + 1u) * 1.52587890625e-05f;
}
static inline float uniform_open(uint32_t w) { return (float)(2u * (w >> 9) + 1u) * 5.9604644775390625e-08f; }

static inline uint32_t draw_word(uint64_t seed, uint32_t task, uint32_t gen, uint32_t purpose,
                                 uint32_t index, uint32_t sub, uint32_t word) {
    uint32_t out[4];
    or_draw_block(seed, task, gen, purpose, index, sub, word >> 2, out);
    return out[word & 3];
}

/* ---- evaluation with the owned, transposed rotation -------------------- */
#define ACC_A(j) ((((j) >> 2) & 1) == 0)

static void eval_points(const task_state* s, int n, const float* x, float* f, float* work) {
    const or_task* t = &s->t;
    const int D = t->dim;
    float* w = work;
    float* z = work + D;
    for (int i = 0; i < n; ++i) {
        const float* xi = x + (size_t)i * D;
        for (int d = 0; d < D; ++d) w[d] = xi[d] - s->shift[d];
        if (t->xform == OR_XFORM_ROTATE) {
            for (int j = 0; j < D; ++j) z[j] = 0.0f;
            for (int d = 0; d < D; ++d) {
                const float wd = w[d];
                const float* q = s->rotT + (size_t)d * D;
                for (int j = 0; j < D; ++j) z[j] = fmaf(q[j], wd, z[j]);
            }
            if (D & 1)
                for (int j = 0; j < D; ++j) z[j] = fmaf(0.0f, 0.0f, z[j]);
        } else if (t->xform == OR_XFORM_DIAG) {
            for (int j = 0; j < D; ++j) z[j] = s->diag[j] * w[j];
        } else {
            for (int j = 0; j < D; ++j) z[j] = w[j];
        }
        f[i] = or_finish32(t, z, s->tab);
    }
}

/* ---- ordering by (fitness, index), NaN last ----------------------------- */
/* rank by counting (thread safe, O(n^2), n <= a few hundred) */
static void order_of(const float* f, int n, int* order) {
    for (int i = 0; i < n; ++i) {
        float fi = f[i];
        if (fi != fi) fi = INFINITY;
        int r = 0;
        for (int p = 0; p < n; ++p) {
            float fp = f[p];
            if (fp != fp) fp = INFINITY;
            if (fp < fi || (fp == fi && p < i)) ++r;
        }
        order[r] = i;
    }
}

/* ---- init ---------------------------------------------------------------- */
static void init_task(or_engine* e, int ti) {
    task_state* s = &e->ts[ti];
    const int D = s->t.dim, P = s->t.pop;
    const float lo = s->t.lo, rng = s->t.hi - s->t.lo;
    for (int i = 0; i < P; ++i) {
        for (int d = 0; d < D; ++d) {
            uint32_t w = draw_word(e->p.seed, (uint32_t)ti, 0u, OR_P_INIT, (uint32_t)i, 0u, (uint32_t)d);
            float u = uniform01(w);
            float t1 = u * rng;
            s->X[(size_t)i * D + d] = t1 + lo;
        }
    }
    float* work = (float*)malloc(sizeof(float) * 2 * (size_t)D);
    eval_points(s, P, s->X, s->fit, work);
    free(work);
}

/* ---- DE index draws -------------------------------------------------------- */
typedef struct { uint32_t blk[4]; int word; int block; uint64_t seed; uint32_t task, gen, index; } idx_stream;

/* word sequence: block0.w0..w2, block1.w0..w3, block2.w0..w3, block3.w0..w3 (15 words) */
static int idx_next(idx_stream* st, uint32_t* out) {
    if (st->block == 0 && st->word == 3) { st->block = 1; st->word = 0;
        or_draw_block(st->seed, st->task, st->gen, OR_P_DE_IDX, st->index, 0u, 1u, st->blk); }
    else if (st->word == 4) {
        if (st->block == 3) return 0;
        st->block += 1; st->word = 0;
        or_draw_block(st->seed, st->task, st->gen, OR_P_DE_IDX, st->index, 0u, (uint32_t)st->block, st->blk);
    }
    *out = st->blk[st->word++];
    return 1;
}

static int draw_excl(idx_stream* st, int m, const int* excl, int nex) {
    int distinct = 0;
    for (int a = 0; a < nex; ++a) {
        int dup = 0;
        for (int b = 0; b < a; ++b) dup |= (excl[b] == excl[a]);
        distinct += !dup;
    }
    for (;;) {
        uint32_t w;
        if (!idx_next(st, &w)) break;
        int r = (int)(w % (uint32_t)m);
        int bad = 0;
        if (distinct < m)
            for (int a = 0; a < nex; ++a) bad |= (excl[a] == r);
        if (!bad) return r;
    }
    /* stream exhausted: smallest index not excluded */
    for (int r = 0; r < m; ++r) {
        int bad = 0;
        for (int a = 0; a < nex; ++a) bad |= (excl[a] == r);
        if (!bad) return r;
    }
    return 0;
}

static inline float clampf(float v, float lo, float hi) {
    if (v > hi) v = hi;
    else if (v < lo) v = lo;
    return v;
}

static void variation_de(or_engine* e, int ti, int gen) {
    task_state* s = &e->ts[ti];
    const int D = s->t.dim, P = s->t.pop, m = P / e->p.n_islands;
    const float lo = s->t.lo, hi = s->t.hi, F = e->p.F, CR = e->p.CR;
    const int rand1 = (e->p.algo == OR_ALGO_DE_RAND_1_BIN);
    for (int i = 0; i < P; ++i) {
        const int il = i % m, base = i - il;
        idx_stream st;
        st.seed = e->p.seed; st.task = (uint32_t)ti; st.gen = (uint32_t)gen; st.index = (uint32_t)i;
        st.block = 0; st.word = 0;
        or_draw_block(st.seed, st.task, st.gen, OR_P_DE_IDX, st.index, 0u, 0u, st.blk);
        const int jrand = (int)(st.blk[3] % (uint32_t)D);
        int ex[4], nex = 0, r1, r2, r3;
        if (rand1) ex[nex++] = il;
        r1 = draw_excl(&st, m, ex, nex); ex[nex++] = r1;
        r2 = draw_excl(&st, m, ex, nex); ex[nex++] = r2;
        r3 = draw_excl(&st, m, ex, nex);
        const float* xb = s->X + (size_t)(base + (rand1 ? r1 : il)) * D;
        const float* x2 = s->X + (size_t)(base + r2) * D;
        const float* x3 = s->X + (size_t)(base + r3) * D;
        const float* xi = s->X + (size_t)i * D;
        float* tr = s->trial + (size_t)i * D;
        /* crossover uniforms: parity h, slot s = d >> 1 uses the 16-bit half (s & 1) of
           word (s >> 1) of sub-stream h (block s >> 3, word (s >> 1) & 3) */
        uint32_t cr[2][4];
        int cached_blk[2] = {-1, -1};
        for (int d = 0; d < D; ++d) {
            float diff = x2[d] - x3[d];
            float sc = F * diff;
            float v = xb[d] + sc;
            v = clampf(v, lo, hi);
            const int h = d & 1, sidx = d >> 1, blk = sidx >> 3;
            if (cached_blk[h] != blk) {
                or_draw_block(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_DE_CR, (uint32_t)i, (uint32_t)h,
                              (uint32_t)blk, cr[h]);
                cached_blk[h] = blk;
            }
            float u = uniform16(cr[h][(sidx >> 1) & 3], sidx & 1);
            tr[d] = (u > CR && d != jrand) ? xi[d] : v;
        }
    }
}

static void variation_ga(or_engine* e, int ti, int gen) {
    task_state* s = &e->ts[ti];
    const int D = s->t.dim, P = s->t.pop, m = P / e->p.n_islands;
    const float lo = s->t.lo, hi = s->t.hi, rng = hi - lo;
    const float ec = 1.0f / (e->p.eta_c + 1.0f);
    const float em1 = e->p.eta_m + 1.0f;
    const float em = 1.0f / em1;
    for (int i = 0; i < P; ++i) {
        const int il = i % m, base = i - il;
        uint32_t mb[4];
        or_draw_block(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_GA_MATE, (uint32_t)(i & ~1), 0u, 0u, mb);
        const int ra = (int)(mb[il & 1] % (uint32_t)m);
        const int rb = (int)(mb[(il & 1) ^ 1] % (uint32_t)m);
        const float* xa = s->X + (size_t)(base + ra) * D;
        const float* xb = s->X + (size_t)(base + rb) * D;
        float* tr = s->trial + (size_t)i * D;
        uint32_t sb[2][4], pb[2][4];
        int cs[2] = {-1, -1}, cp[2] = {-1, -1};
        for (int d = 0; d < D; ++d) {
            const int h = d & 1, sidx = d >> 1;
            if (cs[h] != (sidx >> 2)) {
                or_draw_block(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_GA_SBX, (uint32_t)i, (uint32_t)h,
                              (uint32_t)(sidx >> 2), sb[h]);
                cs[h] = sidx >> 2;
            }
            float p = uniform_open(sb[h][sidx & 3]);
            float tv = (p <= 0.5f) ? (2.0f * p) : (0.5f / (1.0f - p));
            float beta = or_pow(tv, ec);
            if (il & 1) beta = -beta;
            float t1 = 1.0f + beta;
            float t2 = t1 * xa[d];
            float t3 = 1.0f - beta;
            float t4 = t3 * xb[d];
            float t5 = t2 + t4;
            float c = 0.5f * t5;
            c = clampf(c, lo, hi);
            /* polynomial mutation */
            if (cp[h] != (sidx >> 1)) {
                or_draw_block(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_GA_PM, (uint32_t)i, (uint32_t)h,
                              (uint32_t)(sidx >> 1), pb[h]);
                cp[h] = sidx >> 1;
            }
            float r = uniform01(pb[h][(sidx & 1) * 2]);
            if (r < e->p.pm) {
                float u = uniform01(pb[h][(sidx & 1) * 2 + 1]);
                float delta = 1.0f, iv1, iv2, dl;
                float num = c - lo;
                dl = num / rng;
                iv1 = 2.0f * u;
                iv2 = 1.0f - iv1;
                if (u > 0.5f) {
                    delta = -1.0f;
                    float num2 = hi - c;
                    dl = num2 / rng;
                    float omu = 1.0f - u;
                    iv1 = 2.0f * omu;
                    float um = u - 0.5f;
                    iv2 = 2.0f * um;
                }
                float omd = 1.0f - dl;
                float pw = or_pow(omd, em1);
                float a2 = iv2 * pw;
                float a3 = iv1 + a2;
                float a4 = or_pow(a3, em);
                float a5 = a4 - 1.0f;
                delta = delta * a5;
                float step = delta * rng;
                c = c + step;
                c = clampf(c, lo, hi);
            }
            tr[d] = c;
        }
    }
}

static void select_task(task_state* s) {
    const int D = s->t.dim, P = s->t.pop;
    for (int i = 0; i < P; ++i) {
        if (s->fit[i] >= s->tfit[i]) {
            s->fit[i] = s->tfit[i];
            memcpy(s->X + (size_t)i * D, s->trial + (size_t)i * D, sizeof(float) * (size_t)D);
        }
    }
}

static void island_exchange(or_engine* e, int ti, int gen) {
    task_state* s = &e->ts[ti];
    const int D = s->t.dim, P = s->t.pop, n = e->p.n_islands, m = P / n;
    /* shuffled copy into trial/tfit */
    int* arr = s->order;
    for (int q = 0; q < n; ++q) {
        for (int p = 0; p < m; ++p) arr[p] = p;
        for (int k = m - 1; k >= 0; --k) {
            uint32_t w = draw_word(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_ISLX, (uint32_t)q, 0u,
                                   (uint32_t)(m - 1 - k));
            int t = (int)(w % (uint32_t)(k + 1));
            int tmp = arr[k]; arr[k] = arr[t]; arr[t] = tmp;
        }
        for (int p = 0; p < m; ++p) {
            int src = q * m + p, dst = q * m + arr[p];
            memcpy(s->trial + (size_t)dst * D, s->X + (size_t)src * D, sizeof(float) * (size_t)D);
            s->tfit[dst] = s->fit[src];
        }
    }
    /* one island offset per slot k, shared by every island (the reference's
       select_interval[iter * IPB + k] has no block term, SG-EMTF.cu:1227, :1597): slot k of
       island q is refilled from slot k of island (q + off_k) mod n, a rotation, so the
       exchange permutes the task's rows */
    for (int q = 0; q < n; ++q) {
        for (int k = 0; k < m; ++k) {
            uint32_t w = draw_word(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_ISLX, MUT_XCH_INDEX(q), 1u,
                                   (uint32_t)k);
            int src_isl = (q + (int)(w % (uint32_t)n)) % n;
            int src = src_isl * m + k, dst = q * m + k;
            memcpy(s->X + (size_t)dst * D, s->trial + (size_t)src * D, sizeof(float) * (size_t)D);
            s->fit[dst] = s->tfit[src];
        }
    }
}

static void island_migration(or_engine* e, int ti) {
    task_state* s = &e->ts[ti];
    const int D = s->t.dim, P = s->t.pop, n = e->p.n_islands, m = P / n, k = e->p.migrate_count;
    memcpy(s->snapX, s->X, sizeof(float) * (size_t)P * D);
    memcpy(s->snapF, s->fit, sizeof(float) * (size_t)P);
    for (int q = 0; q < n; ++q) order_of(s->snapF + q * m, m, s->order + q * m);
    for (int q = 0; q < n; ++q) {
        const int pdst = (q + 1) % n;
        for (int j = 0; j < k; ++j) {
            int src = q * m + s->order[q * m + MUT_MIG_SRC(m, j)];
            int dst = pdst * m + s->order[pdst * m + MUT_MIG_DST(m, k, j)];
            if (MUT_MIG_BETTER(s->snapF[src], s->snapF[dst])) {
                memcpy(s->X + (size_t)dst * D, s->snapX + (size_t)src * D, sizeof(float) * (size_t)D);
                s->fit[dst] = s->snapF[src];
            }
        }
    }
}

/* Gradient polish of the top-k elites: Armijo steps along -grad32 from a Polyak
 * step f / |g|^2, halving up to 24 times; a point stops at its first failed step. */
#define POLISH_MAX_HALVINGS 24
static void polish_task(or_engine* e, int ti, float* work) {
    task_state* s = &e->ts[ti];
    const or_task* t = &s->t;
    const int D = t->dim, P = t->pop;
    const int K = e->p.polish_topk < P ? e->p.polish_topk : P;
    order_of(s->fit, P, s->order);
    float* x = s->snapX;          /* scratch rows: x, g, xn */
    float* g = x + D;
    float* xn = g + D;
    for (int k = 0; k < K; ++k) {
        const int i = s->order[k];
        memcpy(x, s->X + (size_t)i * D, sizeof(float) * (size_t)D);
        float f = s->fit[i];
        for (int step = 0; step < e->p.polish_steps; ++step) {
            or_grad32(t, 1, x, g);
            float g2 = 0.0f;
            for (int d = 0; d < D; ++d) g2 = fmaf(g[d], g[d], g2);
            if (!(g2 > 0.0f) || !(f > 0.0f)) break;
            float alpha = f / g2;
            int accepted = 0;
            float fn = 0.0f;
            for (int b = 0; b < POLISH_MAX_HALVINGS; ++b) {
                for (int d = 0; d < D; ++d) {
                    float st = alpha * g[d];
                    xn[d] = clampf(x[d] - st, t->lo, t->hi);
                }
                eval_points(s, 1, xn, &fn, work);
                float c1 = 1.0e-4f * alpha;
                float c2 = c1 * g2;
                float thr = f - c2;
                if (fn <= thr) { accepted = 1; break; }
                alpha = alpha * 0.5f;
            }
            if (!accepted) break;
            memcpy(x, xn, sizeof(float) * (size_t)D);
            f = fn;
        }
        memcpy(s->X + (size_t)i * D, x, sizeof(float) * (size_t)D);
        s->fit[i] = f;
    }
}

static void generation(or_engine* e, int ti, int gen, float* work) {
    task_state* s = &e->ts[ti];
    const or_params* p = &e->p;
    if (p->n_islands > 1 && p->exchange_interval > 0 && gen % p->exchange_interval == 0)
        island_exchange(e, ti, gen);
    if (p->algo == OR_ALGO_GA_SBX_PM) variation_ga(e, ti, gen);
    else variation_de(e, ti, gen);
    eval_points(s, s->t.pop, s->trial, s->tfit, work);
    select_task(s);
    if (p->n_islands > 1 && p->migrate_interval > 0 && p->migrate_count > 0 && gen % p->migrate_interval == 0)
        island_migration(e, ti);
    if (p->polish_interval > 0 && p->polish_steps > 0 && p->polish_topk > 0 && gen % p->polish_interval == 0)
        polish_task(e, ti, work);
}

/* ---- inter-task transfer -------------------------------------------------- */
static void snapshot_elites(or_engine* e, int ti) {
    task_state* s = &e->ts[ti];
    const int D = s->t.dim, k = e->p.transfer_count;
    order_of(s->fit, s->t.pop, s->order);
    for (int j = 0; j < k; ++j)
        memcpy(s->elites + (size_t)j * D, s->X + (size_t)s->order[j] * D, sizeof(float) * (size_t)D);
}

/* The donors of the transfer boundary g: the tasks that ran up to it and run on after it
 * (every task of an engine whose tasks all started at 0), ascending ids; returns their count. */
static int donor_list(const or_engine* e, int g, int* out) {
    int nd = 0;
    for (int ti = 0; ti < e->T; ++ti)
        if (MUT_DONOR_BASE(e->ts[ti].base, g) && MUT_DONOR_END(e->ts[ti].end, g)) out[nd++] = ti;
    return nd;
}

/* The index, in donors[0 .. nd), of the donor that the task at position pos draws with w. */
static int donor_pick(uint32_t w, int nd, int pos) {
    int kk = (int)(w % (uint32_t)(nd - 1));
    return kk + (kk >= pos ? 1 : 0);
}

/* Transfer into task ti at its generation gen from a donor drawn among the nd tasks of
 * donors[] (ascending ids, ti among them at position pos): the k-th of the others, k = w mod
 * (nd - 1) -- the reference draws rand() % size over its admitted tasks, again while it hits
 * itself or a task that has finished (SG-EMTF.cu:1696-1699). */
static void transfer_task(or_engine* e, int ti, int gen, const int* donors, int nd, int pos, float* work) {
    task_state* s = &e->ts[ti];
    const int D = s->t.dim, P = s->t.pop, k = e->p.transfer_count;
    uint32_t w = draw_word(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_MIG, 0u, 0u, 0u);
    const int dn = donors[donor_pick(w, nd, pos)];
    const task_state* ds = &e->ts[dn];
    const int Dd = ds->t.dim;
    const int same_box = (ds->t.lo == s->t.lo && ds->t.hi == s->t.hi);
    const float rng_s = s->t.hi - s->t.lo, rng_d = ds->t.hi - ds->t.lo;
    /* self order (snapshot_elites left it in s->order, before any replacement) */
    float* cand = s->trial;
    for (int j = 0; j < k; ++j) {
        const float* self = s->X + (size_t)s->order[j] * D;
        const float* dv = ds->elites + (size_t)j * Dd;
        for (int d = 0; d < D; ++d) {
            float g = self[d];
            if (d < Dd) {
                int take = 1;
                if (e->p.transfer_cr > 0.0f) {
                    uint32_t wu = draw_word(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_XFER, (uint32_t)j,
                                            (uint32_t)(d & 1), (uint32_t)(d >> 1));
                    take = uniform01(wu) > e->p.transfer_cr;
                }
                if (take) {
                    float v = dv[d];
                    if (same_box) g = v;
                    else {
                        float t1 = v - ds->t.lo;
                        float t2 = t1 / MUT_XFER_RD(rng_d, rng_s);
                        float t3 = t2 * MUT_XFER_RS(rng_d, rng_s);
                        g = s->t.lo + t3;
                    }
                }
            }
            cand[(size_t)j * D + d] = g;
        }
    }
    eval_points(s, k, cand, s->tfit, work);
    for (int j = 0; j < k; ++j) {
        int tgt = s->order[MUT_XFER_TGT(P, k, j)];
        if (s->tfit[j] < s->fit[tgt]) {
            s->fit[tgt] = s->tfit[j];
            memcpy(s->X + (size_t)tgt * D, cand + (size_t)j * D, sizeof(float) * (size_t)D);
        }
    }
}

/* ---- public ----------------------------------------------------------------- */
static float* dupf(const float* src, size_t n) {
    float* d = (float*)malloc(sizeof(float) * (n ? n : 1));
    if (src) memcpy(d, src, sizeof(float) * n);
    else memset(d, 0, sizeof(float) * (n ? n : 1));
    return d;
}

static or_engine* create(const or_task* tasks, int32_t n_tasks, const or_params* p, int deferred) {
    or_engine* e = (or_engine*)calloc(1, sizeof(or_engine));
    e->T = n_tasks;
    e->donors = (int*)calloc((size_t)n_tasks, sizeof(int));
    e->p = *p;
    if (e->p.n_islands < 1) e->p.n_islands = 1;
    e->gen = 0;
    e->ts = (task_state*)calloc((size_t)n_tasks, sizeof(task_state));
    for (int ti = 0; ti < n_tasks; ++ti) {
        task_state* s = &e->ts[ti];
        const int D = tasks[ti].dim, P = tasks[ti].pop;
        s->t = tasks[ti];
        s->shift = dupf(tasks[ti].shift, (size_t)D);
        s->diag = dupf(tasks[ti].diag, (size_t)D);
        s->rot = dupf(tasks[ti].rot, tasks[ti].rot ? (size_t)D * D : 0);
        s->rotT = dupf(NULL, (size_t)D * D);
        if (tasks[ti].rot)
            for (int j = 0; j < D; ++j)
                for (int d = 0; d < D; ++d) s->rotT[(size_t)d * D + j] = s->rot[(size_t)j * D + d];
        s->t.rot = s->rot; s->t.shift = s->shift; s->t.diag = s->diag;
        s->tab = dupf(NULL, (size_t)D);
        or_func_table(s->t.func, D, s->tab);
        s->X = dupf(NULL, (size_t)P * D);
        s->fit = dupf(NULL, (size_t)P);
        int kk = P > e->p.transfer_count ? P : e->p.transfer_count;
        s->trial = dupf(NULL, (size_t)kk * D);
        s->tfit = dupf(NULL, (size_t)kk);
        s->snapX = dupf(NULL, (size_t)(P > 3 ? P : 3) * D);
        s->snapF = dupf(NULL, (size_t)P);
        s->order = (int*)calloc((size_t)P, sizeof(int));
        s->elites = dupf(NULL, (size_t)(e->p.transfer_count > 0 ? e->p.transfer_count : 1) * D);
        s->base = deferred ? OR_NEVER : 0;
        s->end = OR_NEVER;
        if (deferred) /* read back before admission: zero rows, NaN fitness (all bits set) */
            memset(s->fit, 0xFF, sizeof(float) * (size_t)P);
    }
    if (!deferred) {
#pragma omp parallel for schedule(dynamic)
        for (int ti = 0; ti < n_tasks; ++ti) init_task(e, ti);
    }
    return e;
}

or_engine* or_create(const or_task* tasks, int32_t n_tasks, const or_params* p) { return create(tasks, n_tasks, p, 0); }

/* Online admission: no task runs until or_admit starts it. */
or_engine* or_create_deferred(const or_task* tasks, int32_t n_tasks, const or_params* p) {
    return create(tasks, n_tasks, p, 1);
}

/* Start tasks first .. first + count - 1 at the current clock for `generations` generations
 * (their initial population is drawn and evaluated now).  Returns 0, or -1 if a task was
 * admitted before, the range is bad, or the clock is not on a transfer boundary. */
int32_t or_admit(or_engine* e, int32_t first, int32_t count, int32_t generations) {
    const int TI = e->p.transfer_interval;
    if (first < 0 || count < 1 || first + count > e->T || generations < 0) return -1;
    if (TI > 0 && e->gen % TI != 0) return -1;
    for (int ti = first; ti < first + count; ++ti)
        if (e->ts[ti].base != OR_NEVER) return -1;
    for (int ti = first; ti < first + count; ++ti) {
        e->ts[ti].base = e->gen;
        e->ts[ti].end = e->gen + generations;
    }
#pragma omp parallel for schedule(dynamic)
    for (int ti = first; ti < first + count; ++ti) init_task(e, ti);
    return 0;
}

void or_destroy(or_engine* e) {
    if (!e) return;
    for (int ti = 0; ti < e->T; ++ti) {
        task_state* s = &e->ts[ti];
        free(s->rot); free(s->rotT); free(s->shift); free(s->diag); free(s->tab); free(s->X); free(s->fit);
        free(s->trial); free(s->tfit); free(s->snapX); free(s->snapF); free(s->order); free(s->elites);
    }
    free(e->ts);
    free(e->donors);
    free(e);
}

void or_run(or_engine* e, int32_t n_gens, int32_t n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    const int g_end = e->gen + n_gens;
    const int TI = e->p.transfer_interval;
    const int xfer_on = (TI > 0 && e->p.transfer_count > 0);
    while (e->gen < g_end) {
        const int g = e->gen;
        if (xfer_on && g % TI == 0) {
            const int nd = donor_list(e, g, e->donors);
            if (nd >= 2) {
#pragma omp parallel for schedule(dynamic)
                for (int q = 0; q < nd; ++q) snapshot_elites(e, e->donors[q]);
#pragma omp parallel for schedule(dynamic)
                for (int q = 0; q < nd; ++q) {
                    const int ti = e->donors[q];
                    float* work = (float*)malloc(sizeof(float) * 2 * (size_t)e->ts[ti].t.dim);
                    transfer_task(e, ti, MUT_LOCAL_GEN(g, e->ts[ti].base), e->donors, nd, MUT_DONOR_POS(q, ti), work);
                    free(work);
                }
            }
        }
        /* epochs end at transfer boundaries and where a running task ends */
        int g_stop = g_end;
        if (xfer_on) {
            int nb = (g / TI + 1) * TI;
            if (nb < g_stop) g_stop = nb;
        }
        for (int ti = 0; ti < e->T; ++ti)
            if (e->ts[ti].base <= g && e->ts[ti].end > g && e->ts[ti].end < g_stop) g_stop = e->ts[ti].end;
#pragma omp parallel for schedule(dynamic)
        for (int ti = 0; ti < e->T; ++ti) {
            task_state* s = &e->ts[ti];
            if (!(s->base <= g && s->end >= g_stop)) continue;
            float* work = (float*)malloc(sizeof(float) * 2 * (size_t)s->t.dim);
            for (int gg = g + 1; gg <= g_stop; ++gg) generation(e, ti, MUT_LOCAL_GEN(gg, s->base), work);
            free(work);
        }
        e->gen = g_stop;
    }
}

int32_t or_generation(const or_engine* e) { return e->gen; }

/* The donor task `task` draws at the transfer boundary `clock` under the current admissions
 * (-1 if it receives nothing there). */
int32_t or_transfer_donor(const or_engine* e, int32_t task, int32_t clock) {
    const int TI = e->p.transfer_interval;
    if (TI <= 0 || e->p.transfer_count <= 0 || clock % TI != 0 || task < 0 || task >= e->T) return -1;
    int* list = (int*)malloc(sizeof(int) * (size_t)e->T);
    const int nd = donor_list(e, clock, list);
    int pos = -1;
    for (int q = 0; q < nd; ++q)
        if (list[q] == task) pos = q;
    int dn = -1;
    if (pos >= 0 && nd >= 2) {
        const uint32_t w = draw_word(e->p.seed, (uint32_t)task, (uint32_t)MUT_LOCAL_GEN(clock, e->ts[task].base),
                                     OR_P_MIG, 0u, 0u, 0u);
        dn = list[donor_pick(w, nd, MUT_DONOR_POS(pos, task))];
    }
    free(list);
    return dn;
}

/* Single phases on the current state, for the tests that pin them one at a time. */
void or_phase_exchange(or_engine* e, int32_t task, int32_t gen) {
    if (e->p.n_islands > 1) island_exchange(e, task, gen);
}

void or_phase_migrate(or_engine* e, int32_t task) {
    if (e->p.n_islands > 1 && e->p.migrate_count > 0) island_migration(e, task);
}

void or_phase_transfer(or_engine* e, int32_t gen) {
    if (e->T < 2 || e->p.transfer_count < 1) return;
    for (int ti = 0; ti < e->T; ++ti) e->donors[ti] = ti;
    for (int ti = 0; ti < e->T; ++ti) snapshot_elites(e, ti);
    for (int ti = 0; ti < e->T; ++ti) {
        float* work = (float*)malloc(sizeof(float) * 2 * (size_t)e->ts[ti].t.dim);
        transfer_task(e, ti, gen, e->donors, e->T, ti, work);
        free(work);
    }
}

void or_get_population(const or_engine* e, int32_t task, float* x, float* fit) {
    const task_state* s = &e->ts[task];
    if (x) memcpy(x, s->X, sizeof(float) * (size_t)s->t.pop * s->t.dim);
    if (fit) memcpy(fit, s->fit, sizeof(float) * (size_t)s->t.pop);
}

void or_set_population(or_engine* e, int32_t task, const float* x, const float* fit) {
    task_state* s = &e->ts[task];
    if (x) memcpy(s->X, x, sizeof(float) * (size_t)s->t.pop * s->t.dim);
    if (fit) memcpy(s->fit, fit, sizeof(float) * (size_t)s->t.pop);
}

void or_get_best(const or_engine* e, float* fit, float* x_best, int32_t x_stride) {
    for (int ti = 0; ti < e->T; ++ti) {
        const task_state* s = &e->ts[ti];
        int best = 0;
        for (int i = 1; i < s->t.pop; ++i) {
            float fb = s->fit[best], fi = s->fit[i];
            if (fb != fb) fb = INFINITY;
            if (fi != fi) fi = INFINITY;
            if (fi < fb) best = i;
        }
        if (fit) fit[ti] = s->fit[best];
        if (x_best) memcpy(x_best + (size_t)ti * x_stride, s->X + (size_t)best * s->t.dim, sizeof(float) * (size_t)s->t.dim);
    }
}
