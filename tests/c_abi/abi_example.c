/* abi_example.c -- a plain C99 caller of the engine's C ABI (include/emtf.h), the way a
 * reference-side host program would drive it: 16 shifted Sphere / Rastrigin tasks,
 * D = 20, pop 64, DE/rand/1/bin with elite transfer, 200 generations, with the best of
 * every task recorded on the device every 50 generations.  Prints one line
 * "ok <best of task 0> <best of task 1> <trace points> <task 0 at the last point>
 * <second-best of task 0>" or the library's error message.  Built and run by
 * tests/test_abi.py. */
#include <stdio.h>
#include <stdlib.h>
#include "emtf.h"

int main(void) {
    enum { T = 16, D = 20, POP = 64 };
    static float shift[T][D];
    emtf_task_desc tasks[T];
    for (int t = 0; t < T; ++t) {
        for (int d = 0; d < D; ++d) shift[t][d] = (float)((t * 7 + d * 3) % 41) - 20.0f;
        tasks[t].func = (t % 2) ? EMTF_RASTRIGIN : EMTF_SPHERE;
        tasks[t].dim = D;
        tasks[t].pop = POP;
        tasks[t].xform = EMTF_SHIFT;
        tasks[t].lo = -100.0f;
        tasks[t].hi = 100.0f;
        tasks[t].rot = NULL;
        tasks[t].shift = shift[t];
        tasks[t].diag = NULL;
    }
    emtf_params p = {0};
    p.algo = EMTF_DE_RAND_1_BIN;
    p.F = 0.5f;
    p.CR = 0.9f;
    p.n_islands = 1;
    p.transfer_interval = 20;
    p.transfer_count = 2;
    p.seed = 7;
    emtf_handle* h = NULL;
    if (emtf_create(tasks, T, &p, &h) != EMTF_OK) {
        printf("error %s\n", emtf_last_error());
        return 1;
    }
    float best[T];
    static float xb[T][D];
    static float trace[5][T], worst[T];
    static float ef[T][2], ex[T][2][D];
    int32_t points = 0;
    if (emtf_trace_start(h, 50, 5, NULL) != EMTF_OK || emtf_run(h, 200, NULL) != EMTF_OK ||
        emtf_sync(h) != EMTF_OK || emtf_get_best(h, best, &xb[0][0], D) != EMTF_OK ||
        emtf_trace(h, &points, &trace[0][0], worst) != EMTF_OK ||
        emtf_get_elites(h, 2, &ef[0][0], &ex[0][0][0], D) != EMTF_OK) {
        printf("error %s\n", emtf_last_error());
        emtf_destroy(h);
        return 1;
    }
    if (points != 5 || trace[4][0] != best[0] || ef[0][0] != best[0]) {
        printf("error inconsistent exports\n");
        emtf_destroy(h);
        return 1;
    }
    printf("ok %.9g %.9g %d %.9g %.9g\n", best[0], best[1], (int)points, trace[4][0], ef[0][1]);
    emtf_destroy(h);
    return 0;
}
