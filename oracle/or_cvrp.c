/*
 * or_cvrp.c -- capacitated VRP genetic algorithm of the oracle (exact integers).
 * TEST INFRASTRUCTURE ONLY (see emtf_oracle.h); also timed as the CPU baseline.
 *
 * The reference's CVRP variant (proposed model/SG-EMTF_cvrp.cu) runs the same
 * island engine on permutation-like float genomes, but its operators are a
 * non-functional prototype (crossover writes all-ones genomes, :731-736; the
 * init shuffle never runs, :171).  This restates the intended algorithm with the
 * textbook components the benchmark configuration names:
 *   distances  TSPLIB EUC_2D: d(i,j) = nint(sqrt(dx^2 + dy^2)), integers
 *   genome     giant tour: a permutation of customers 1..n (depot 0 implicit)
 *   decode     Prins (2004) split: shortest path over the tour, a route per arc
 *              (i..j-1) with load <= capacity, cost d(0,t_i) + path + d(t_{j-1},0)
 *   crossover  Davis order crossover (OX), cut points a <= b
 *   mutation   swap of two positions, or 2-opt (reverse a segment)
 *   selection  child replaces its parent iff cost(child) <= cost(parent)
 *              (the reference's greedy rule, SG-EMTF_cvrp.cu Selection)
 *   migration  intra-instance island migration, as the continuous engine
 * with Philox draws addressed like the continuous engine (purposes 16, 17).
 */
#include "emtf_oracle.h"
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CV_RING 64

/* Planted mistake for the mutation tests (see or_engine.c): OR_MUTANT 7 sends rank j to
 * destination rank m-1-j.  The oracle is built with OR_MUTANT undefined (0). */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif
#define MUT_CV_DST(m, k, j) (OR_MUTANT == 7 ? (m) - 1 - (j) : (m) - (k) + (j))

typedef struct {
    int n, cap, pop;
    uint8_t* dist;     /* (n+1) x (n+1) */
    uint8_t* dem;      /* n+1 */
    uint8_t* P;        /* pop x n */
    uint8_t* C;        /* pop x n children */
    int32_t* cost;     /* pop */
    int32_t* ccost;    /* pop */
    int* order;        /* pop */
} cv_state;

struct or_cvrp {
    int T, gen;
    or_cvrp_params p;
    cv_state* s;
};

int32_t or_cvrp_dist(int32_t x0, int32_t y0, int32_t x1, int32_t y1) {
    const double dx = (double)(x0 - x1), dy = (double)(y0 - y1);
    return (int32_t)(sqrt(dx * dx + dy * dy) + 0.5);
}

static inline uint32_t wordof(uint64_t seed, uint32_t task, uint32_t gen, uint32_t purpose, uint32_t index,
                              uint32_t sub, uint32_t w) {
    uint32_t b[4];
    or_draw_block(seed, task, gen, purpose, index, sub, w >> 2, b);
    return b[w & 3];
}

/* Prins split: minimal cost of cutting the giant tour t[0..n) into feasible routes.
 * Pull form over a ring of the last CV_RING labels (routes hold < CV_RING customers). */
int32_t or_cvrp_split(const uint8_t* t, int n, const uint8_t* dist, const uint8_t* dem, int cap) {
    int32_t V[CV_RING];
    const int S = n + 1;
    V[0] = 0;
    for (int j = 1; j <= n; ++j) {
        int32_t best = INT32_MAX;
        int load = 0, path = 0;
        const int last = t[j - 1];
        const int back = dist[last * S + 0];
        for (int i = j - 1; i >= 0; --i) {
            load += dem[t[i]];
            if (load > cap) break;
            if (i < j - 1) path += dist[t[i] * S + t[i + 1]];
            const int32_t c = V[i & (CV_RING - 1)] + dist[0 * S + t[i]] + path + back;
            if (c < best) best = c;
        }
        V[j & (CV_RING - 1)] = best;
    }
    return V[n & (CV_RING - 1)];
}

/* child of p1 and p2: OX with cuts a <= b, then swap / 2-opt mutation */
void or_cvrp_child(const uint8_t* p1, const uint8_t* p2, int n, int a, int b, int mut, int c, int d,
                   uint8_t* child) {
    uint32_t used[8] = {0};
    for (int k = a; k <= b; ++k) {
        child[k] = p1[k];
        used[p1[k] >> 5] |= 1u << (p1[k] & 31);
    }
    int k = (b + 1) % n;
    for (int s = 0; s < n; ++s) {
        const int g = p2[(b + 1 + s) % n];
        if (used[g >> 5] & (1u << (g & 31))) continue;
        child[k] = (uint8_t)g;
        k = (k + 1) % n;
    }
    if (mut == 1) {
        const uint8_t tmp = child[c];
        child[c] = child[d];
        child[d] = tmp;
    } else if (mut == 2) {
        int lo = c < d ? c : d, hi = c < d ? d : c;
        while (lo < hi) {
            const uint8_t tmp = child[lo];
            child[lo] = child[hi];
            child[hi] = tmp;
            ++lo;
            --hi;
        }
    }
}

static void cv_rank(const int32_t* f, int n, int* order) {
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int p = 0; p < n; ++p) r += (f[p] < f[i] || (f[p] == f[i] && p < i));
        order[r] = i;
    }
}

/* Island migration on a snapshot of the costs: rank j of island q replaces rank m-k+j of
 * island q+1 (mod n) iff strictly cheaper, as the continuous engine's island_migration. */
static void cv_migrate(or_cvrp* e, int ti) {
    cv_state* s = &e->s[ti];
    const int n = s->n, P = s->pop, nis = e->p.n_islands, m = P / nis, k = e->p.migrate_count;
    for (int q = 0; q < nis; ++q) cv_rank(s->cost + q * m, m, s->order + q * m);
    memcpy(s->ccost, s->cost, sizeof(int32_t) * (size_t)P);
    for (int q = 0; q < nis; ++q) {
        const int pd = (q + 1) % nis;
        for (int j = 0; j < k; ++j) {
            const int src = q * m + s->order[q * m + j], dst = pd * m + s->order[pd * m + MUT_CV_DST(m, k, j)];
            if (s->ccost[src] < s->ccost[dst]) {
                s->cost[dst] = s->ccost[src];
                memcpy(s->P + (size_t)dst * n, s->P + (size_t)src * n, (size_t)n);
            }
        }
    }
}

static void cv_generation(or_cvrp* e, int ti, int gen) {
    cv_state* s = &e->s[ti];
    const int n = s->n, P = s->pop, m = P / e->p.n_islands;
    const float thr1 = e->p.p_swap, thr2 = e->p.p_swap + e->p.p_2opt;
    for (int i = 0; i < P; ++i) {
        const int il = i % m, base = i - il;
        uint32_t b0[4], b1[4];
        or_draw_block(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_CV_GEN, (uint32_t)i, 0u, 0u, b0);
        or_draw_block(e->p.seed, (uint32_t)ti, (uint32_t)gen, OR_P_CV_GEN, (uint32_t)i, 1u, 0u, b1);
        int r = (int)(b0[0] % (uint32_t)(m - 1));
        if (r >= il) r += 1;
        int a = (int)(b0[1] % (uint32_t)n), b = (int)(b0[2] % (uint32_t)n);
        if (a > b) { const int t = a; a = b; b = t; }
        const float u = (float)((b0[3] >> 8) + 1u) * 5.9604644775390625e-08f;
        const int mut = (u < thr1) ? 1 : ((u < thr2) ? 2 : 0);
        const int c = (int)(b1[0] % (uint32_t)n), d = (int)(b1[1] % (uint32_t)n);
        uint8_t* child = s->C + (size_t)i * n;
        or_cvrp_child(s->P + (size_t)i * n, s->P + (size_t)(base + r) * n, n, a, b, mut, c, d, child);
        s->ccost[i] = or_cvrp_split(child, n, s->dist, s->dem, s->cap);
    }
    for (int i = 0; i < P; ++i) {
        if (s->ccost[i] <= s->cost[i]) {
            s->cost[i] = s->ccost[i];
            memcpy(s->P + (size_t)i * n, s->C + (size_t)i * n, (size_t)n);
        }
    }
    if (e->p.n_islands > 1 && e->p.migrate_interval > 0 && e->p.migrate_count > 0 &&
        gen % e->p.migrate_interval == 0)
        cv_migrate(e, ti);
}

or_cvrp* or_cvrp_create(const or_cvrp_instance* inst, int32_t n_inst, const or_cvrp_params* p) {
    or_cvrp* e = (or_cvrp*)calloc(1, sizeof(or_cvrp));
    e->T = n_inst;
    e->p = *p;
    if (e->p.n_islands < 1) e->p.n_islands = 1;
    e->s = (cv_state*)calloc((size_t)n_inst, sizeof(cv_state));
    for (int t = 0; t < n_inst; ++t) {
        cv_state* s = &e->s[t];
        const int n = inst[t].n, P = inst[t].pop, S = n + 1;
        s->n = n;
        s->cap = inst[t].capacity;
        s->pop = P;
        s->dist = (uint8_t*)malloc((size_t)S * S);
        s->dem = (uint8_t*)malloc((size_t)S);
        for (int i = 0; i < S; ++i) {
            s->dem[i] = (uint8_t)inst[t].demand[i];
            for (int j = 0; j < S; ++j)
                s->dist[i * S + j] = (uint8_t)or_cvrp_dist(inst[t].xy[2 * i], inst[t].xy[2 * i + 1],
                                                           inst[t].xy[2 * j], inst[t].xy[2 * j + 1]);
        }
        s->P = (uint8_t*)malloc((size_t)P * n);
        s->C = (uint8_t*)malloc((size_t)P * n);
        s->cost = (int32_t*)malloc(sizeof(int32_t) * (size_t)P);
        s->ccost = (int32_t*)malloc(sizeof(int32_t) * (size_t)P);
        s->order = (int*)malloc(sizeof(int) * (size_t)P);
    }
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < n_inst; ++t) {
        cv_state* s = &e->s[t];
        const int n = s->n;
        for (int i = 0; i < s->pop; ++i) {
            uint8_t* tour = s->P + (size_t)i * n;
            for (int k = 0; k < n; ++k) tour[k] = (uint8_t)(k + 1);
            for (int k = n - 1; k >= 1; --k) {
                const uint32_t w = wordof(e->p.seed, (uint32_t)t, 0u, OR_P_CV_INIT, (uint32_t)i, 0u,
                                          (uint32_t)(n - 1 - k));
                const int j = (int)(w % (uint32_t)(k + 1));
                const uint8_t tmp = tour[k];
                tour[k] = tour[j];
                tour[j] = tmp;
            }
            s->cost[i] = or_cvrp_split(tour, n, s->dist, s->dem, s->cap);
        }
    }
    return e;
}

void or_cvrp_run(or_cvrp* e, int32_t gens, int32_t n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    const int g0 = e->gen;
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < e->T; ++t)
        for (int g = g0 + 1; g <= g0 + gens; ++g) cv_generation(e, t, g);
    e->gen = g0 + gens;
}

void or_cvrp_get(const or_cvrp* e, int32_t inst, uint8_t* tours, int32_t* cost) {
    const cv_state* s = &e->s[inst];
    if (tours) memcpy(tours, s->P, (size_t)s->pop * s->n);
    if (cost) memcpy(cost, s->cost, sizeof(int32_t) * (size_t)s->pop);
}

void or_cvrp_phase_migrate(or_cvrp* e, int32_t inst) {
    if (e->p.n_islands > 1 && e->p.migrate_count > 0) cv_migrate(e, inst);
}

void or_cvrp_set(or_cvrp* e, int32_t inst, const uint8_t* tours, const int32_t* cost) {
    cv_state* s = &e->s[inst];
    if (tours) memcpy(s->P, tours, (size_t)s->pop * s->n);
    if (cost) memcpy(s->cost, cost, sizeof(int32_t) * (size_t)s->pop);
}

void or_cvrp_destroy(or_cvrp* e) {
    if (!e) return;
    for (int t = 0; t < e->T; ++t) {
        cv_state* s = &e->s[t];
        free(s->dist); free(s->dem); free(s->P); free(s->C); free(s->cost); free(s->ccost); free(s->order);
    }
    free(e->s);
    free(e);
}
