// cvrp.hip -- capacitated-VRP genetic algorithm on gfx950 (exact integer arithmetic).
//
// One workgroup owns one instance for an epoch of generations; one lane owns one
// individual.  Everything an epoch touches lives in LDS: the giant tours (uint8
// customer ids), the children, the node coordinates (relative to the depot, int16
// pairs), the depot distances and demands, and each lane's ring of split values.
// Distances between customers are computed when needed: nint(sqrt(dx^2 + dy^2)) of
// integer differences is exact in binary32 (sqrt of an integer <= 65280 lies at least
// 4.9e-4 from every half-integer, far above the square root's rounding), and keeping
// no (n+1)^2 matrix lets two instances of 200 customers share a CU.  Per generation a lane
//   1. draws its mate, OX cut points and mutation from Philox (registers),
//   2. builds its child with Davis order crossover (OX) + swap / 2-opt mutation,
//   3. decodes it with the Prins split (shortest path over the giant tour,
//      routes of load <= capacity) in int32, as a sliding-window minimum over a
//      per-lane ring of packed (label term, prefix demand) entries,
//   4. keeps the child iff its cost is <= the parent's (reference Selection rule,
//      proposed model/SG-EMTF_cvrp.cu).
// The reference's CVRP kernel (SG-EMTF_cvrp.cu:668-962) is a prototype whose
// crossover overwrites genomes with ones; this implements the algorithm it names.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "emtf_math.h"
#include "cvrp_types.h"

namespace emtf {

constexpr uint32_t P_CV_INIT = 16, P_CV_GEN = 17;
constexpr int kRing = 64;  // ring positions kept per lane; routes hold < 64 customers
constexpr int kMirror = 16; // slots 0 .. 15 repeated at 64 .. 79, so any 16 consecutive positions are contiguous

// Lane stride of the split ring in entries: at least kRing + kMirror, and an odd number of
// 4-byte words, so the 64 lanes of a wave reading the same slot hit 64 different banks.
template <typename RingT>
__host__ __device__ constexpr int ring_stride() {
    return sizeof(RingT) == 2 ? 82 : 81;
}

// RingT: uint16_t entries (W + woff) when every instance's W + woff range fits 16 bits,
// else int32_t (W itself, woff = 0).
template <typename RingT>
struct CvLds {
    uint8_t* P;     // popP x NP tours
    uint8_t* C;     // popP x NP children
    uint32_t* xy;   // S node coordinates relative to the depot, (x & 0xFFFF) | (y << 16)
    uint8_t* d0;    // DS depot distances d(0, c)
    uint8_t* dem;   // DS demands
    RingT* R;       // popP x ring_stride split ring (lane-major)
    uint32_t* used; // 8 x popP OX bitmaps
    int32_t* cost;  // popP
    int32_t* ccost; // popP
    int32_t* ord;   // popP
    int32_t* aux;   // popP
};

__host__ __device__ inline int cvrp_lds_bytes(int n, int popP, int ring_bytes) {
    const int NP = (n + 3) & ~3, DS = (n + 1 + 3) & ~3;
    const int ring = ring_bytes == 2 ? ring_stride<uint16_t>() * 2 : ring_stride<int32_t>() * 4;
    return popP * ring + 8 * popP * 4 + 4 * popP * 4 + 4 * DS + 2 * DS + 2 * popP * NP;
}

// TSPLIB EUC_2D distance nint(sqrt(dx^2 + dy^2)) of two packed relative coordinates.
// Every distance of an instance is <= 255 (checked at create), so s = dx^2 + dy^2 < 255.5^2,
// and sqrt(s) of an integer s lies at least 0.25 / 511 = 4.9e-4 from every half-integer;
// the hardware square root (v_sqrt_f32, 1 ulp <= 2^-16 below 256) and the rounding of
// the + 0.5 stay far inside that, so the rounded distance is exact without the
// correctly rounded square root's refinement.
__device__ __forceinline__ int nint_xy(uint32_t a, uint32_t b) {
    const int dx = (int)(int16_t)(a & 0xFFFFu) - (int)(int16_t)(b & 0xFFFFu);
    const int dy = (int)(int16_t)(a >> 16) - (int)(int16_t)(b >> 16);
    const float s = (float)(dx * dx + dy * dy);
    return (int)(__builtin_amdgcn_sqrtf(s) + 0.5f);
}

// Prins split of the lane's tour t (LDS row) as a sliding-window minimum.  With path
// prefix sums P_k = sum_{m<k} d(t_m, t_{m+1}), the route t_i .. t_{j-1} costs
// d(0, t_i) + P_{j-1} - P_i + d(t_{j-1}, 0), so the label of j is
//   V_j = P_{j-1} + d(t_{j-1}, 0) + min W_i,  W_i = V_i + d(0, t_i) - P_i,
// over the route starts i in [lo, j) whose demand fits the capacity (V_0 = 0) -- the same
// integer sums and minimum as the pull form, so the same labels.  The window start lo
// follows j (the demand of t_lo .. t_{j-1} kept in a register).  The lane's ring holds
// W_i + woff for the last kRing positions (position i in slot i mod 64, and slots 0 .. 15
// again in 64 .. 79, so the newest 16 positions sit in 16 consecutive slots read at
// immediate offsets from one address); the slots of positions -16 .. -1 start at the
// largest value and every entry is overwritten with it when it leaves the window, so the
// minimum over the newest 16 positions needs no mask (older slots hold the largest value
// or, aliased by the ring, a value of the window) -- windows longer than 16 continue in
// blocks of 8.
template <typename RingT>
__device__ __forceinline__ int32_t split_cost(const uint8_t* t, int n, const CvLds<RingT>& L, int cap, int lane,
                                              int popP, int woff) {
    constexpr RingT NONE = (RingT)((sizeof(RingT) == 2) ? 0xFFFF : 0x7FFFFFFF);
    constexpr int kS = ring_stride<RingT>();
    RingT* R = L.R + lane * kS;
    // every position >= 0 is written (value, or the largest value once it leaves the
    // window) before any read of it, and reads reach back to position -16 only (the
    // newest 16 from j = 1 on; the blocks from lo - 7 >= -7): slots 48 .. 63 hold the
    // negative positions, and they alone start at the largest value
#pragma unroll
    for (int k = kRing - 16; k < kRing; ++k) R[k] = NONE;
    // writes slot s and its mirror (the same slot again when s >= 16)
    auto put = [&](int pos, RingT v) {
        const int s = pos & (kRing - 1);
        R[s] = v;
        R[s < kMirror ? s + kRing : s] = v;
    };
    const int cur = t[0];
    uint32_t xyc = L.xy[cur];
    int d0c = L.d0[cur];   // d(0, t_{j-1}) = d(t_{j-1}, 0)
    int dcur = L.dem[cur];
    int32_t P = 0;         // P_{j-1}
    int lo = 0, wload = 0; // window start, demand of t_lo .. t_{j-1}
    // loaded one step or one window move ahead, so neither waits on its own load: the
    // customer t_j, the demands of t_lo and t_{lo+1}, the customer t_{lo+2} (clamped to the
    // tour; a clamped one is never reached, since lo <= n - 1)
    int nxt = n > 1 ? t[1] : 0;
    int dlo = dcur, dlo1 = n > 1 ? L.dem[t[1]] : 0;
    int tl2 = t[n > 2 ? 2 : n - 1];
    put(0, (RingT)(d0c + woff));  // W_0 = d(0, t_0)
    int32_t V = 0;
    for (int j = 1; j <= n; ++j) {
        // the next customer's coordinates, distances and demand do not depend on this label
        const int nn = (j + 1 < n) ? t[j + 1] : 0;
        const uint32_t xyn = L.xy[nxt];
        const int d0n = L.d0[nxt];
        const int dmn = L.dem[nxt];
        const int dnx = nint_xy(xyc, xyn);
        wload += dcur;
        while (wload > cap) {  // lo <= j - 1: every demand fits the capacity
            put(lo, NONE);
            wload -= dlo;
            dlo = dlo1;
            dlo1 = L.dem[tl2];
            ++lo;
            tl2 = t[lo + 2 < n ? lo + 2 : n - 1];
        }
        // positions j - 16 .. j - 1: slots s0 .. s0 + 15 with s0 = (j - 16) mod 64
        const RingT* Rw = R + ((j - 16) & (kRing - 1));
        int32_t w[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) w[k] = (int32_t)Rw[k];
#pragma unroll
        for (int h = 8; h >= 1; h >>= 1) {
#pragma unroll
            for (int k = 0; k < h; ++k) w[k] = w[k + h] < w[k] ? w[k + h] : w[k];
        }
        int32_t m = w[0];
        for (int i0 = j - 17; i0 >= lo; i0 -= 8) {
            // positions i0 - 7 .. i0: slots s1 .. s1 + 7 with s1 = (i0 - 7) mod 64
            const RingT* Rb = R + ((i0 - 7) & (kRing - 1));
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int32_t v = (int32_t)Rb[k];
                m = v < m ? v : m;
            }
        }
        V = P + d0c + (m - woff);
        if (j == n) break;
        P += dnx;
        put(j, (RingT)(V + d0n - P + woff));
        nxt = nn;
        xyc = xyn;
        d0c = d0n;
        dcur = dmn;
    }
    return V;
}

template <typename RingT>
__device__ __forceinline__ void make_child(const uint8_t* p1, const uint8_t* p2, uint8_t* child, int n, int a, int b,
                                           int mut, int c, int d, const CvLds<RingT>& L, int lane, int popP) {
    uint32_t* used = L.used + lane;
#pragma unroll
    for (int w = 0; w < 8; ++w) used[w * popP] = 0u;
    for (int k = a; k <= b; ++k) {
        const int g = p1[k];
        child[k] = (uint8_t)g;
        used[(g >> 5) * popP] |= 1u << (g & 31);
    }
    // the rest of p2 from b + 1 on, in order, skipping the genes taken from p1: four
    // genes and their bitmap words per round of loads
    int k = (b + 1 == n) ? 0 : b + 1;
    int src = k;
    for (int s = 0; s < n; s += 4) {
        int g[4];
        uint32_t u[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int sq = src + q;           // < n + 3
            const int iq = sq < n ? sq : sq - n;
            g[q] = p2[iq < n ? iq : 0];       // n < 4: a gene past the tour is read in the row, unused
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) u[q] = used[(g[q] >> 5) * popP];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (s + q < n && !((u[q] >> (g[q] & 31)) & 1u)) {
                child[k] = (uint8_t)g[q];
                k = (k + 1 == n) ? 0 : k + 1;
            }
        }
        src = src + 4 < n ? src + 4 : src + 4 - n;
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

template <typename RingT>
__device__ __forceinline__ void rank_cost(const CvLds<RingT>& L, int pop, int m, int tid, int nthreads) {
    for (int i = tid; i < pop; i += nthreads) {
        const int q = i / m, base = q * m, il = i - base;
        const int32_t fi = L.cost[i];
        int r = 0;
        for (int p = 0; p < m; ++p) {
            const int32_t fp = L.cost[base + p];
            r += (fp < fi || (fp == fi && p < il)) ? 1 : 0;
        }
        L.ord[base + r] = i;
    }
}

template <typename RingT>
__global__ __launch_bounds__(256) void cvrp_epoch_kernel(CvrpArgs a, const int* __restrict__ order) {
    extern __shared__ __align__(16) uint8_t cv_smem[];
    const int inst = order[blockIdx.x];
    const CvrpDev cd = a.inst[inst];
    const int n = cd.n, cap = cd.cap, pop = cd.pop, S = n + 1;
    const int NP = (n + 3) & ~3, DS = (n + 1 + 3) & ~3;
    const int popP = blockDim.x, tid = threadIdx.x;
    CvLds<RingT> L;
    uint8_t* p = cv_smem;
    L.R = reinterpret_cast<RingT*>(p);
    p += ring_stride<RingT>() * popP * sizeof(RingT);
    L.used = reinterpret_cast<uint32_t*>(p);
    p += 8 * popP * 4;
    L.cost = reinterpret_cast<int32_t*>(p);
    L.ccost = L.cost + popP;
    L.ord = L.ccost + popP;
    L.aux = L.ord + popP;
    p += 4 * popP * 4;
    L.xy = reinterpret_cast<uint32_t*>(p);
    p += 4 * DS;
    L.d0 = p;
    p += DS;
    L.dem = p;
    p += DS;
    L.P = p;
    p += popP * NP;
    L.C = p;
    const Rng rng{a.seed_lo, a.seed_hi, (uint32_t)inst};
    const int m = pop / (a.n_islands > 0 ? a.n_islands : 1);

    for (int e = tid; e < S; e += popP) {
        const uint32_t c = a.xy[cd.xy_off + e];
        L.xy[e] = c;
        L.d0[e] = (uint8_t)nint_xy(0u, c);
        L.dem[e] = a.dem[cd.dem_off + e];
    }
    uint8_t* tg = a.tours + cd.tour_off;
    if (!a.init) {
        for (int e = tid; e < pop * n; e += popP) {
            const int i = e / n, k = e - i * n;
            L.P[i * NP + k] = tg[e];
        }
        for (int i = tid; i < pop; i += popP) L.cost[i] = a.cost[cd.cost_off + i];
    }
    __syncthreads();
    if (a.init && tid < pop) {
        uint8_t* tour = L.P + tid * NP;
        for (int k = 0; k < n; ++k) tour[k] = (uint8_t)(k + 1);
        U4 blk{0u, 0u, 0u, 0u};
        for (int k = n - 1; k >= 1; --k) {
            const int w = n - 1 - k;
            if ((w & 3) == 0) blk = rng.block(0u, P_CV_INIT, (uint32_t)tid, 0u, (uint32_t)(w >> 2));
            const int j = (int)(pick(blk, (uint32_t)(w & 3)) % (uint32_t)(k + 1));
            const uint8_t tmp = tour[k];
            tour[k] = tour[j];
            tour[j] = tmp;
        }
        L.cost[tid] = split_cost(tour, n, L, cap, tid, popP, cd.woff);
    }
    __syncthreads();

    const float thr1 = a.p_swap, thr2 = a.p_swap + a.p_2opt;
    for (int gen = a.g0 + 1; gen <= a.g1; ++gen) {
        if (tid < pop) {
            const int il = tid % m, base = tid - il;
            const U4 b0 = rng.block((uint32_t)gen, P_CV_GEN, (uint32_t)tid, 0u, 0u);
            const U4 b1 = rng.block((uint32_t)gen, P_CV_GEN, (uint32_t)tid, 1u, 0u);
            int r = (int)(b0.x % (uint32_t)(m - 1));
            if (r >= il) r += 1;
            int ca = (int)(b0.y % (uint32_t)n), cb = (int)(b0.z % (uint32_t)n);
            if (ca > cb) { const int t = ca; ca = cb; cb = t; }
            const float u = uniform01(b0.w);
            const int mut = (u < thr1) ? 1 : ((u < thr2) ? 2 : 0);
            const int c = (int)(b1.x % (uint32_t)n), d = (int)(b1.y % (uint32_t)n);
            uint8_t* child = L.C + tid * NP;
            make_child(L.P + tid * NP, L.P + (base + r) * NP, child, n, ca, cb, mut, c, d, L, tid, popP);
            L.ccost[tid] = split_cost(child, n, L, cap, tid, popP, cd.woff);
        }
        __syncthreads();
        if (tid < pop && L.ccost[tid] <= L.cost[tid]) {
            L.cost[tid] = L.ccost[tid];
            uint32_t* dst = reinterpret_cast<uint32_t*>(L.P + tid * NP);
            const uint32_t* src = reinterpret_cast<const uint32_t*>(L.C + tid * NP);
            for (int w = 0; w < NP / 4; ++w) dst[w] = src[w];
        }
        __syncthreads();
        const int nis = a.n_islands, kk = a.migrate_count;
        if (nis > 1 && a.migrate_interval > 0 && kk > 0 && gen % a.migrate_interval == 0) {
            rank_cost(L, pop, m, tid, popP);
            __syncthreads();
            for (int e = tid; e < nis * kk; e += popP) {
                const int q = e / kk, j = e - q * kk, pd = (q + 1) % nis;
                const int src = L.ord[q * m + j], dst = L.ord[pd * m + m - kk + j];
                L.aux[e] = (L.cost[src] < L.cost[dst]) ? 1 : 0;
            }
            __syncthreads();
            for (int e = tid; e < nis * kk * (NP / 4); e += popP) {
                const int pr = e / (NP / 4), w = e - pr * (NP / 4);
                if (L.aux[pr]) {
                    const int q = pr / kk, j = pr - q * kk, pd = (q + 1) % nis;
                    const int src = L.ord[q * m + j], dst = L.ord[pd * m + m - kk + j];
                    reinterpret_cast<uint32_t*>(L.P + dst * NP)[w] = reinterpret_cast<const uint32_t*>(L.P + src * NP)[w];
                }
            }
            for (int e = tid; e < nis * kk; e += popP) {
                if (L.aux[e]) {
                    const int q = e / kk, j = e - q * kk, pd = (q + 1) % nis;
                    const int src = L.ord[q * m + j], dst = L.ord[pd * m + m - kk + j];
                    L.cost[dst] = L.cost[src];
                }
            }
            __syncthreads();
        }
    }
    for (int e = tid; e < pop * n; e += popP) {
        const int i = e / n, k = e - i * n;
        tg[e] = L.P[i * NP + k];
    }
    for (int i = tid; i < pop; i += popP) a.cost[cd.cost_off + i] = L.cost[i];
}

}  // namespace emtf

extern "C" int emtf_cvrp_lds_bytes(int n, int popP, int ring_bytes) {
    return emtf::cvrp_lds_bytes(n, popP, ring_bytes);
}

// ring_bytes 2: W + woff in 16 bits (the host checks that every instance's W + woff range
// fits); else W in 32 bits.
extern "C" int emtf_cvrp_launch(const emtf::CvrpArgs* args, const int* order, int n_inst, int threads, int lds_bytes,
                                int ring_bytes, hipStream_t stream) {
    static bool lds_ok[2] = {false, false};
    const int w = ring_bytes == 2 ? 0 : 1;
    const void* k = w == 0 ? reinterpret_cast<const void*>(&emtf::cvrp_epoch_kernel<uint16_t>)
                           : reinterpret_cast<const void*>(&emtf::cvrp_epoch_kernel<int32_t>);
    if (!lds_ok[w]) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return (int)e;
        lds_ok[w] = true;
    }
    if (w == 0)
        hipLaunchKernelGGL(emtf::cvrp_epoch_kernel<uint16_t>, dim3(n_inst), dim3(threads), lds_bytes, stream, *args, order);
    else
        hipLaunchKernelGGL(emtf::cvrp_epoch_kernel<int32_t>, dim3(n_inst), dim3(threads), lds_bytes, stream, *args, order);
    return (int)hipGetLastError();
}
