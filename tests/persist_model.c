/* CPU model of the persistent schedule's ordering protocol (test infrastructure).
 *
 * Host threads stand in for workgroups (fewer threads than items, so not every item is
 * held by a running worker at once).  Items (epoch e, task) are taken e-major from one
 * atomic counter.  An item waits for its task's previous epoch (stamp[task] >= e) and,
 * when it transfers, for its donor's previous epoch (stamp[donor] >= e); it reads the
 * donor's elites from mailbox buffer e & 1, then counts itself in reads[e]; after its
 * work, an item that publishes first waits for reads[e - 1] == n_tasks (every reader of
 * the buffer it is about to overwrite is past its read), writes its tag into buffer
 * (e + 1) & 1, and releases stamp[task] = e + 1.  Every wait is bounded; a wait that runs
 * out sets the error flag and every worker drains.
 *
 * Checked: every item runs once, after its task's previous epoch; every mailbox read sees
 * exactly the donor's publication of the previous epoch, before and after a random pause
 * in the middle of the read (an overwrite during the read shows); no wait runs out.
 *
 * usage: persist_model n_tasks n_epochs n_threads seed [mistake]
 *   mistake 1: no donor wait; 2: no wait for the readers before publishing;
 *   3: no wait for the task's own previous epoch (each must be caught for some seed).
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <unistd.h>

static int T, E, NTH, MISTAKE;
static uint64_t SEED;
static atomic_int next_item, err_flag, violations;
static atomic_int *stamp, *reads;
static atomic_int *mailbox[2];
static int *state, *order, *donor; /* donor[e * T + t] */
static atomic_int *runs;           /* per item */

static uint64_t mix(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
    return x;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* up to max_us, and one time in 32 a long stall of 2 - 10 ms (a slow workgroup) */
static void pause_us(uint64_t r, int max_us) {
    int us = (int)(r % (uint64_t)(max_us + 1));
    if (((r >> 40) & 31) == 0) us = 2000 + (int)((r >> 20) % 8001);
    if (us > 0) usleep(us);
}

static int tag(int t, int e) { return (e + 1) * 100000 + t; } /* e = -1: the previous launch */

/* bounded wait for *p >= v; 0 when it runs out or another worker's wait did */
static int wait_at_least(atomic_int* p, int v) {
    const double t0 = now_s();
    for (;;) {
        if (atomic_load_explicit(p, memory_order_acquire) >= v) return 1;
        if (atomic_load_explicit(&err_flag, memory_order_relaxed)) return 0;
        if (now_s() - t0 > 5.0) {
            atomic_store(&err_flag, 1);
            return 0;
        }
        sched_yield();
    }
}

static void violation(const char* what, int e, int t) {
    if (atomic_fetch_add(&violations, 1) < 8) fprintf(stderr, "violation: %s (epoch %d task %d)\n", what, e, t);
}

static void* worker(void* arg) {
    uint64_t r = mix(SEED * 1000003ULL + (uint64_t)(intptr_t)arg);
    for (;;) {
        const int item = atomic_fetch_add(&next_item, 1);
        if (item >= T * E) break;
        const int e = item / T, t = order[item - e * T];
        atomic_fetch_add(&runs[e * T + t], 1);
        r = mix(r);
        int ok = 1;
        if (MISTAKE != 3) ok = wait_at_least(&stamp[t], e);
        const int dn = donor[e * T + t];
        if (ok && MISTAKE != 1) ok = wait_at_least(&stamp[dn], e);
        if (!ok) continue;
        if (state[t] != e) violation("task's previous epoch not finished", e, t);
        /* the transfer read, with a pause in the middle */
        const int v1 = atomic_load_explicit(&mailbox[e & 1][dn], memory_order_relaxed);
        pause_us(r, 30);
        r = mix(r);
        const int v2 = atomic_load_explicit(&mailbox[e & 1][dn], memory_order_relaxed);
        if (v1 != tag(dn, e - 1) || v2 != tag(dn, e - 1)) violation("mailbox read not the donor's previous epoch", e, t);
        atomic_fetch_add_explicit(&reads[e], 1, memory_order_release);
        /* the epoch's work */
        pause_us(r, 200);
        r = mix(r);
        if (e > 0 && MISTAKE != 2) ok = wait_at_least(&reads[e - 1], T);
        if (!ok) continue;
        atomic_store_explicit(&mailbox[(e + 1) & 1][t], tag(t, e), memory_order_relaxed);
        state[t] = e + 1;
        atomic_store_explicit(&stamp[t], e + 1, memory_order_release);
    }
    return NULL;
}

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s n_tasks n_epochs n_threads seed [mistake]\n", argv[0]);
        return 2;
    }
    T = atoi(argv[1]);
    E = atoi(argv[2]);
    NTH = atoi(argv[3]);
    SEED = (uint64_t)strtoull(argv[4], NULL, 10);
    MISTAKE = argc > 5 ? atoi(argv[5]) : 0;
    if (T < 2 || E < 1 || NTH < 1) return 2;
    stamp = calloc(T, sizeof *stamp);
    reads = calloc(E, sizeof *reads);
    mailbox[0] = calloc(T, sizeof **mailbox);
    mailbox[1] = calloc(T, sizeof **mailbox);
    state = calloc(T, sizeof *state);
    order = calloc(T, sizeof *order);
    donor = calloc((size_t)T * E, sizeof *donor);
    runs = calloc((size_t)T * E, sizeof *runs);
    uint64_t r = mix(SEED + 17);
    for (int t = 0; t < T; ++t) order[t] = t;
    for (int t = T - 1; t > 0; --t) { /* a fixed task order, as the cost order is */
        r = mix(r);
        const int j = (int)(r % (uint64_t)(t + 1));
        const int x = order[t]; order[t] = order[j]; order[j] = x;
    }
    for (int e = 0; e < E; ++e)
        for (int t = 0; t < T; ++t) {
            r = mix(r);
            int d = (int)(r % (uint64_t)(T - 1));
            donor[e * T + t] = d >= t ? d + 1 : d;
        }
    /* the previous launch published epoch -1 into buffer 0 */
    for (int t = 0; t < T; ++t) atomic_store(&mailbox[0][t], tag(t, -1));
    pthread_t* th = calloc(NTH, sizeof *th);
    for (int i = 0; i < NTH; ++i) pthread_create(&th[i], NULL, worker, (void*)(intptr_t)i);
    for (int i = 0; i < NTH; ++i) pthread_join(th[i], NULL);
    int bad = atomic_load(&violations);
    if (atomic_load(&err_flag)) {
        fprintf(stderr, "a wait ran out\n");
        ++bad;
    }
    for (int i = 0; i < T * E; ++i)
        if (atomic_load(&runs[i]) != 1) {
            fprintf(stderr, "item %d ran %d times\n", i, atomic_load(&runs[i]));
            ++bad;
            break;
        }
    if (!atomic_load(&err_flag))
        for (int t = 0; t < T; ++t)
            if (state[t] != E) { fprintf(stderr, "task %d finished %d epochs\n", t, state[t]); ++bad; break; }
    printf(bad ? "FAIL\n" : "OK\n");
    return bad ? 1 : 0;
}
