// cvrp_host.cpp -- host runtime of the CVRP genetic algorithm behind include/emtf.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/emtf.h"
#include "cvrp_types.h"

extern "C" int emtf_cvrp_launch(const emtf::CvrpArgs* args, const int* order, int n_inst, int threads, int lds_bytes,
                                int ring_bytes, hipStream_t stream);
extern "C" int emtf_cvrp_lds_bytes(int n, int popP, int ring_bytes);
extern "C" void emtf_set_error(const char* msg);

namespace {
int cfail(int code, const std::string& msg) {
    emtf_set_error(msg.c_str());
    return code;
}
constexpr int kMaxLds = 160 * 1024;
}  // namespace

// Instances are launched in groups of equal occupancy (workgroups per CU allowed by
// their LDS footprint), each group with its own LDS size on its own stream, forked from
// and joined back into the caller's stream per epoch.
struct CvGroup {
    int offset = 0, count = 0, lds = 0, stream = 0;
};

struct emtf_cvrp_handle {
    int T = 0, gen = 0, threads = 64, ring_bytes = 2;  // split ring entry size (2 or 4)
    std::vector<CvGroup> groups;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev_join;
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_order = nullptr;  // orders a run on a new stream after the previous one
    emtf::CvrpArgs a{};
    std::vector<emtf::CvrpDev> inst;
    emtf::CvrpDev* d_inst = nullptr;
    int* d_order = nullptr;
    uint8_t* d_tours = nullptr;
    int32_t* d_cost = nullptr;
    uint32_t* d_xy = nullptr;
    uint8_t* d_dem = nullptr;
    hipStream_t stream = nullptr;
};

static void cv_free(emtf_cvrp_handle* h) {
    if (!h) return;
    (void)hipFree(h->d_inst);
    (void)hipFree(h->d_order);
    (void)hipFree(h->d_tours);
    (void)hipFree(h->d_cost);
    (void)hipFree(h->d_xy);
    (void)hipFree(h->d_dem);
    for (hipStream_t s : h->side) (void)hipStreamDestroy(s);
    for (hipEvent_t e : h->ev_join) (void)hipEventDestroy(e);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_order) (void)hipEventDestroy(h->ev_order);
    delete h;
}

// TSPLIB EUC_2D distance of integer points: nint(sqrt(dx^2 + dy^2))
static int nint_dist(int x0, int y0, int x1, int y1) {
    const double dx = (double)(x0 - x1), dy = (double)(y0 - y1);
    return (int)(std::sqrt(dx * dx + dy * dy) + 0.5);
}

static int cv_launch(emtf_cvrp_handle* h, int g0, int g1, int init, hipStream_t s) {
    h->a.g0 = g0;
    h->a.g1 = g1;
    h->a.init = init;
    const bool fork = !h->side.empty();
    if (fork) {
        if (hipEventRecord(h->ev_fork, s) != hipSuccess) return cfail(EMTF_E_DEVICE, "hipEventRecord");
        for (hipStream_t q : h->side)
            if (hipStreamWaitEvent(q, h->ev_fork, 0) != hipSuccess) return cfail(EMTF_E_DEVICE, "hipStreamWaitEvent");
    }
    for (const CvGroup& g : h->groups) {
        hipStream_t q = g.stream == 0 ? s : h->side[g.stream - 1];
        const int rc = emtf_cvrp_launch(&h->a, h->d_order + g.offset, g.count, h->threads, g.lds, h->ring_bytes, q);
        if (rc != 0) return cfail(EMTF_E_DEVICE, std::string("cvrp kernel launch: ") + hipGetErrorString((hipError_t)rc));
    }
    if (fork) {
        for (size_t k = 0; k < h->side.size(); ++k)
            if (hipEventRecord(h->ev_join[k], h->side[k]) != hipSuccess || hipStreamWaitEvent(s, h->ev_join[k], 0) != hipSuccess)
                return cfail(EMTF_E_DEVICE, "stream join");
    }
    return EMTF_OK;
}

extern "C" {

int emtf_cvrp_create(const emtf_cvrp_instance* in, int32_t n_inst, const emtf_cvrp_params* prm,
                     emtf_cvrp_handle** out) {
    emtf_set_error("");
    if (!out) return cfail(EMTF_E_INVALID, "out is NULL");
    *out = nullptr;
    if (!in || !prm || n_inst < 1) return cfail(EMTF_E_INVALID, "instances/params missing or n < 1");
    const int nis = prm->n_islands < 1 ? 1 : prm->n_islands;
    if (prm->migrate_interval < 0 || prm->migrate_count < 0) return cfail(EMTF_E_INVALID, "negative migration knob");
    int maxpop = 0;
    for (int t = 0; t < n_inst; ++t) {
        const emtf_cvrp_instance& c = in[t];
        char buf[160];
        if (c.n < 2 || c.n > 255) { snprintf(buf, sizeof buf, "instance %d: n %d outside [2, 255]", t, c.n); return cfail(EMTF_E_INVALID, buf); }
        if (!c.xy || !c.demand) { snprintf(buf, sizeof buf, "instance %d: coordinates/demands missing", t); return cfail(EMTF_E_INVALID, buf); }
        if (c.pop < 2 || c.pop > 256 || c.pop % nis || c.pop / nis < 2) { snprintf(buf, sizeof buf, "instance %d: pop %d (islands %d)", t, c.pop, nis); return cfail(EMTF_E_INVALID, buf); }
        int dmin = 1 << 30, dmax = 0;
        for (int i = 1; i <= c.n; ++i) { dmin = std::min(dmin, c.demand[i]); dmax = std::max(dmax, c.demand[i]); }
        if (dmin < 1 || dmax > 255 || c.demand[0] != 0) { snprintf(buf, sizeof buf, "instance %d: demands must be in [1, 255], depot 0", t); return cfail(EMTF_E_INVALID, buf); }
        if (c.capacity < dmax) { snprintf(buf, sizeof buf, "instance %d: capacity below a demand", t); return cfail(EMTF_E_INVALID, buf); }
        if (c.capacity / dmin >= 64) { snprintf(buf, sizeof buf, "instance %d: routes could exceed 63 customers", t); return cfail(EMTF_E_UNSUPPORTED, buf); }
        if (nis > 1 && prm->migrate_interval > 0 && 2 * prm->migrate_count > c.pop / nis) { snprintf(buf, sizeof buf, "instance %d: migrate_count > island/2", t); return cfail(EMTF_E_INVALID, buf); }
        maxpop = std::max(maxpop, c.pop);
    }
    emtf_cvrp_handle* h = new emtf_cvrp_handle();
    h->T = n_inst;
    h->threads = (maxpop + 63) & ~63;
    h->inst.resize(n_inst);
    int64_t nt = 0, nc = 0, nd = 0, nx = 0, ne = 0;
    for (int t = 0; t < n_inst; ++t) {
        emtf::CvrpDev& d = h->inst[t];
        std::memset(&d, 0, sizeof d);
        d.n = in[t].n;
        d.cap = in[t].capacity;
        d.pop = in[t].pop;
        d.tour_off = nt;
        d.cost_off = nc;
        d.xy_off = nx;
        d.dem_off = ne;
        nt += (int64_t)d.pop * d.n;
        nc += d.pop;
        nd += (int64_t)(d.n + 1) * (d.n + 1);
        nx += d.n + 1;
        ne += d.n + 1;
    }
    // the distance matrix stays on the host (checks and bounds); the kernel gets the
    // coordinates relative to the depot and computes the same nint distances
    std::vector<uint8_t> dist((size_t)nd), dem((size_t)ne);
    std::vector<uint32_t> xyr((size_t)nx);
    for (int64_t t = 0, off = 0; t < n_inst; ++t) {
        const emtf_cvrp_instance& c = in[t];
        const int S = c.n + 1;
        for (int i = 0; i < S; ++i) {
            dem[h->inst[t].dem_off + i] = (uint8_t)c.demand[i];
            const int rx = c.xy[2 * i] - c.xy[0], ry = c.xy[2 * i + 1] - c.xy[1];
            xyr[h->inst[t].xy_off + i] = ((uint32_t)rx & 0xFFFFu) | ((uint32_t)ry << 16);
            for (int j = 0; j < S; ++j) {
                const int dv = nint_dist(c.xy[2 * i], c.xy[2 * i + 1], c.xy[2 * j], c.xy[2 * j + 1]);
                if (dv > 255) { cv_free(h); return cfail(EMTF_E_UNSUPPORTED, "distance above 255"); }
                dist[off + (int64_t)i * S + j] = (uint8_t)dv;
            }
        }
        off += (int64_t)S * S;
    }
    // Split ring entries W_i + woff with W_i = V_i + d(0, t_i) - P_i: a label V_i never
    // exceeds the cost of serving each customer alone, sum_c 2 d(0, c), and a path prefix
    // P_i never exceeds (n - 1) dmax, so woff = (n - 1) dmax makes every entry non-negative
    // and bounds it by woff + sum_c 2 d(0, c) + dmax.  The ring holds 16-bit entries when
    // that bound fits for every instance, else W itself in 32 bits (woff = 0).
    h->ring_bytes = 2;
    std::vector<int64_t> woffs(n_inst);
    for (int64_t t = 0, off = 0; t < n_inst; ++t) {
        const int S = in[t].n + 1;
        int64_t vbound = 0, dmax = 0;
        for (int c = 1; c < S; ++c) vbound += 2 * (int64_t)dist[off + c];
        for (int64_t e = 0; e < (int64_t)S * S; ++e) dmax = std::max<int64_t>(dmax, dist[off + e]);
        woffs[t] = (int64_t)(in[t].n - 1) * dmax;
        if (woffs[t] + vbound + dmax >= 65536) h->ring_bytes = 4;
        off += (int64_t)S * S;
    }
    for (int t = 0; t < n_inst; ++t) h->inst[t].woff = h->ring_bytes == 2 ? (int32_t)woffs[t] : 0;
    // groups of equal occupancy (workgroups per CU from the LDS footprint), longest
    // instances first inside each group
    std::vector<int> occ(n_inst);
    for (int t = 0; t < n_inst; ++t) {
        const int b = emtf_cvrp_lds_bytes(in[t].n, h->threads, h->ring_bytes);
        if (b > kMaxLds) { cv_free(h); return cfail(EMTF_E_UNSUPPORTED, "instance exceeds LDS"); }
        occ[t] = std::min(8, kMaxLds / b);
    }
    std::vector<int> order(n_inst);
    for (int t = 0; t < n_inst; ++t) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (occ[x] != occ[y]) return occ[x] < occ[y];
        return in[x].n * in[x].pop > in[y].n * in[y].pop;
    });
    for (int r = 0; r < n_inst;) {
        CvGroup g;
        g.offset = r;
        while (r < n_inst && occ[order[r]] == occ[order[g.offset]]) {
            g.lds = std::max(g.lds, emtf_cvrp_lds_bytes(in[order[r]].n, h->threads, h->ring_bytes));
            ++r;
        }
        g.count = r - g.offset;
        g.stream = (int)(h->groups.size() % 4);
        h->groups.push_back(g);
    }
    const int n_side = std::min<int>(3, (int)h->groups.size() - 1);
    for (int k = 0; k < n_side; ++k) {
        hipStream_t q = nullptr;
        hipEvent_t ev = nullptr;
        if (hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            cv_free(h);
            return cfail(EMTF_E_DEVICE, "stream / event creation failed");
        }
        h->side.push_back(q);
        h->ev_join.push_back(ev);
    }
    if (n_side > 0 && hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess) {
        cv_free(h);
        return cfail(EMTF_E_DEVICE, "event creation failed");
    }
    auto ok = [&](hipError_t e) { return e == hipSuccess; };
    if (!ok(hipMalloc((void**)&h->d_inst, sizeof(emtf::CvrpDev) * n_inst)) || !ok(hipMalloc((void**)&h->d_order, sizeof(int) * n_inst)) ||
        !ok(hipMalloc((void**)&h->d_tours, (size_t)nt)) || !ok(hipMalloc((void**)&h->d_cost, sizeof(int32_t) * (size_t)nc)) ||
        !ok(hipMalloc((void**)&h->d_xy, sizeof(uint32_t) * (size_t)nx)) || !ok(hipMalloc((void**)&h->d_dem, (size_t)ne))) {
        cv_free(h);
        return cfail(EMTF_E_NOMEM, "hipMalloc failed");
    }
    if (!ok(hipMemcpy(h->d_inst, h->inst.data(), sizeof(emtf::CvrpDev) * n_inst, hipMemcpyHostToDevice)) ||
        !ok(hipMemcpy(h->d_order, order.data(), sizeof(int) * n_inst, hipMemcpyHostToDevice)) ||
        !ok(hipMemcpy(h->d_xy, xyr.data(), sizeof(uint32_t) * (size_t)nx, hipMemcpyHostToDevice)) ||
        !ok(hipMemcpy(h->d_dem, dem.data(), (size_t)ne, hipMemcpyHostToDevice))) {
        cv_free(h);
        return cfail(EMTF_E_DEVICE, "hipMemcpy failed");
    }
    emtf::CvrpArgs& a = h->a;
    a.inst = h->d_inst;
    a.n_inst = n_inst;
    a.tours = h->d_tours;
    a.cost = h->d_cost;
    a.xy = h->d_xy;
    a.dem = h->d_dem;
    a.p_swap = prm->p_swap;
    a.p_2opt = prm->p_2opt;
    a.n_islands = nis;
    a.migrate_interval = prm->migrate_interval;
    a.migrate_count = prm->migrate_count;
    a.seed_lo = (uint32_t)prm->seed;
    a.seed_hi = (uint32_t)(prm->seed >> 32);
    int rc = cv_launch(h, 0, 0, 1, nullptr);
    if (rc == EMTF_OK && hipStreamSynchronize(nullptr) != hipSuccess) rc = cfail(EMTF_E_DEVICE, "cvrp init failed");
    if (rc != EMTF_OK) { cv_free(h); return rc; }
    *out = h;
    return EMTF_OK;
}

int emtf_cvrp_run(emtf_cvrp_handle* h, int32_t generations, void* stream) {
    if (!h) return cfail(EMTF_E_INVALID, "handle is NULL");
    if (generations < 0) return cfail(EMTF_E_INVALID, "generations < 0");
    // a run on another stream than the previous call's waits for that call's work
    if ((hipStream_t)stream != h->stream) {
        if ((!h->ev_order && hipEventCreateWithFlags(&h->ev_order, hipEventDisableTiming) != hipSuccess) ||
            hipEventRecord(h->ev_order, h->stream) != hipSuccess ||
            hipStreamWaitEvent((hipStream_t)stream, h->ev_order, 0) != hipSuccess)
            return cfail(EMTF_E_DEVICE, "stream ordering");
        h->stream = (hipStream_t)stream;
    }
    // epochs of at most 100 generations keep each launch short
    while (generations > 0) {
        const int g = std::min(generations, 100);
        const int rc = cv_launch(h, h->gen, h->gen + g, 0, h->stream);
        if (rc != EMTF_OK) return rc;
        h->gen += g;
        generations -= g;
    }
    return EMTF_OK;
}

int emtf_cvrp_sync(emtf_cvrp_handle* h) {
    if (!h) return cfail(EMTF_E_INVALID, "handle is NULL");
    if (hipStreamSynchronize(h->stream) != hipSuccess) return cfail(EMTF_E_DEVICE, "hipStreamSynchronize");
    return EMTF_OK;
}

int emtf_cvrp_get_population(emtf_cvrp_handle* h, int32_t i, uint8_t* tours, int32_t* cost) {
    if (!h || i < 0 || i >= h->T) return cfail(EMTF_E_INVALID, "bad handle or instance index");
    if (hipStreamSynchronize(h->stream) != hipSuccess) return cfail(EMTF_E_DEVICE, "hipStreamSynchronize");
    const emtf::CvrpDev& d = h->inst[i];
    if (tours && hipMemcpy(tours, h->d_tours + d.tour_off, (size_t)d.pop * d.n, hipMemcpyDeviceToHost) != hipSuccess)
        return cfail(EMTF_E_DEVICE, "hipMemcpy tours");
    if (cost && hipMemcpy(cost, h->d_cost + d.cost_off, sizeof(int32_t) * d.pop, hipMemcpyDeviceToHost) != hipSuccess)
        return cfail(EMTF_E_DEVICE, "hipMemcpy cost");
    return EMTF_OK;
}

int emtf_cvrp_get_best(emtf_cvrp_handle* h, int32_t* cost, uint8_t* tour, int32_t stride) {
    if (!h) return cfail(EMTF_E_INVALID, "handle is NULL");
    std::vector<int32_t> c;
    std::vector<uint8_t> t;
    for (int i = 0; i < h->T; ++i) {
        const emtf::CvrpDev& d = h->inst[i];
        if (tour && stride < d.n) return cfail(EMTF_E_INVALID, "stride < n");
        c.resize(d.pop);
        t.resize((size_t)d.pop * d.n);
        const int rc = emtf_cvrp_get_population(h, i, tour ? t.data() : nullptr, c.data());
        if (rc != EMTF_OK) return rc;
        int best = 0;
        for (int k = 1; k < d.pop; ++k)
            if (c[k] < c[best]) best = k;
        if (cost) cost[i] = c[best];
        if (tour) std::memcpy(tour + (size_t)i * stride, t.data() + (size_t)best * d.n, d.n);
    }
    return EMTF_OK;
}

int emtf_cvrp_destroy(emtf_cvrp_handle* h) {
    if (!h) return EMTF_OK;
    (void)hipStreamSynchronize(h->stream);
    cv_free(h);
    return EMTF_OK;
}

}  // extern "C"
