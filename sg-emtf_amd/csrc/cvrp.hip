// cvrp.hip -- capacitated-VRP genetic algorithm on gfx950 (exact integer arithmetic).
//
// One workgroup owns one instance for an epoch of generations; one lane owns one
// individual.  Everything an epoch touches lives in LDS: the giant tours (uint8
// customer ids), the children, the (n+1)^2 uint8 distance matrix (nint Euclidean
// distances of integer coordinates are <= 141 on [0,100]^2), demands, and each
// lane's ring of split labels.  Per generation a lane
//   1. draws its mate, OX cut points and mutation from Philox (registers),
//   2. builds its child with Davis order crossover (OX) + swap / 2-opt mutation,
//   3. decodes it with the Prins split (shortest path over the giant tour,
//      routes of load <= capacity) in int32,
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
constexpr int kRing = 64;  // split labels kept per lane; routes hold < 64 customers
constexpr int kSplitBlock = 4;  // steps of the split's inner loop per round of loads (4 beat 6 and 8)

template <typename LabelT>
struct CvLds {
    uint8_t* P;     // popP x NP tours
    uint8_t* C;     // popP x NP children
    uint8_t* dist;  // (n+1) x DS
    uint8_t* dem;   // DS
    LabelT* V;      // kRing x popP split labels (uint16 when every label of the batch fits)
    uint32_t* used; // 8 x popP OX bitmaps
    int32_t* cost;  // popP
    int32_t* ccost; // popP
    int32_t* ord;   // popP
    int32_t* aux;   // popP
};

__host__ __device__ inline int cvrp_lds_bytes(int n, int popP, int label_bytes) {
    const int NP = (n + 3) & ~3, DS = (n + 1 + 3) & ~3;
    return 2 * popP * NP + (n + 1) * DS + DS + kRing * popP * label_bytes + 8 * popP * 4 + 4 * popP * 4;
}

// Prins split of the lane's tour t (LDS row); labels in the lane's ring column.
// The pull loop over route starts i = j-1, j-2, ... runs in blocks of kSplitBlock
// steps: the block's customer ids are loaded, then their demands, distances and labels, then
// the steps are applied in order until the load exceeds the capacity -- two rounds
// of LDS latency per block, with the sums and the minimum formed exactly as one
// step at a time.  Loads past the stopping step stay in bounds (position < 0 reads
// the depot; ring slots wrap) and are not used.
template <typename LabelT>
__device__ __forceinline__ int32_t split_cost(const uint8_t* t, int n, const CvLds<LabelT>& L, int S, int cap, int lane,
                                              int popP) {
    L.V[0 * popP + lane] = 0;
    for (int j = 1; j <= n; ++j) {
        int32_t best = 0x7fffffff;
        int load = 0, path = 0;
        const int last = t[j - 1];
        const int back = L.dist[last * S + 0];
        int prev = last;  // customer at position i + 1 (for i = j - 1 its distance is unused)
        bool stop = false;
        for (int i0 = j - 1; i0 >= 0 && !stop; i0 -= kSplitBlock) {
            int tk[kSplitBlock], dm[kSplitBlock], dd[kSplitBlock], d0[kSplitBlock];
            int32_t v[kSplitBlock];
#pragma unroll
            for (int k = 0; k < kSplitBlock; ++k) tk[k] = (i0 - k >= 0) ? t[i0 - k] : 0;
#pragma unroll
            for (int k = 0; k < kSplitBlock; ++k) {
                const int nx = (k == 0) ? prev : tk[k - 1];
                dm[k] = L.dem[tk[k]];
                dd[k] = L.dist[tk[k] * S + nx];
                d0[k] = L.dist[tk[k]];
                v[k] = (int32_t)L.V[((i0 - k) & (kRing - 1)) * popP + lane];
            }
#pragma unroll
            for (int k = 0; k < kSplitBlock; ++k) {
                const int i = i0 - k;
                if (stop || i < 0) break;
                load += dm[k];
                if (load > cap) {
                    stop = true;
                    break;
                }
                if (i < j - 1) path += dd[k];
                const int32_t c = v[k] + d0[k] + path + back;
                best = c < best ? c : best;
            }
            prev = tk[kSplitBlock - 1];
        }
        L.V[(j & (kRing - 1)) * popP + lane] = (LabelT)best;
    }
    return (int32_t)L.V[(n & (kRing - 1)) * popP + lane];
}

template <typename LabelT>
__device__ __forceinline__ void make_child(const uint8_t* p1, const uint8_t* p2, uint8_t* child, int n, int a, int b,
                                           int mut, int c, int d, const CvLds<LabelT>& L, int lane, int popP) {
    uint32_t* used = L.used + lane;
#pragma unroll
    for (int w = 0; w < 8; ++w) used[w * popP] = 0u;
    for (int k = a; k <= b; ++k) {
        const int g = p1[k];
        child[k] = (uint8_t)g;
        used[(g >> 5) * popP] |= 1u << (g & 31);
    }
    int k = (b + 1 == n) ? 0 : b + 1;
    int src = k;
    for (int s = 0; s < n; ++s) {
        const int g = p2[src];
        src = (src + 1 == n) ? 0 : src + 1;
        if (used[(g >> 5) * popP] & (1u << (g & 31))) continue;
        child[k] = (uint8_t)g;
        k = (k + 1 == n) ? 0 : k + 1;
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

template <typename LabelT>
__device__ __forceinline__ void rank_cost(const CvLds<LabelT>& L, int pop, int m, int tid, int nthreads) {
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

template <typename LabelT>
__global__ __launch_bounds__(256) void cvrp_epoch_kernel(CvrpArgs a, const int* __restrict__ order) {
    extern __shared__ __align__(16) uint8_t cv_smem[];
    const int inst = order[blockIdx.x];
    const CvrpDev cd = a.inst[inst];
    const int n = cd.n, cap = cd.cap, pop = cd.pop, S = n + 1;
    const int NP = (n + 3) & ~3, DS = (n + 1 + 3) & ~3;
    const int popP = blockDim.x, tid = threadIdx.x;
    CvLds<LabelT> L;
    uint8_t* p = cv_smem;
    L.V = reinterpret_cast<LabelT*>(p);
    p += kRing * popP * sizeof(LabelT);
    L.used = reinterpret_cast<uint32_t*>(p);
    p += 8 * popP * 4;
    L.cost = reinterpret_cast<int32_t*>(p);
    L.ccost = L.cost + popP;
    L.ord = L.ccost + popP;
    L.aux = L.ord + popP;
    p += 4 * popP * 4;
    L.P = p;
    p += popP * NP;
    L.C = p;
    p += popP * NP;
    L.dem = p;
    p += DS;
    L.dist = p;
    (void)DS;
    const Rng rng{a.seed_lo, a.seed_hi, (uint32_t)inst};
    const int m = pop / (a.n_islands > 0 ? a.n_islands : 1);

    const uint8_t* dg = a.dist + cd.dist_off;
    for (int e = tid; e < S * S; e += popP) {
        const int i = e / S, j = e - i * S;
        L.dist[i * S + j] = dg[e];
    }
    for (int e = tid; e < S; e += popP) L.dem[e] = a.dem[cd.dem_off + e];
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
        L.cost[tid] = split_cost(tour, n, L, S, cap, tid, popP);
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
            L.ccost[tid] = split_cost(child, n, L, S, cap, tid, popP);
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

extern "C" int emtf_cvrp_lds_bytes(int n, int popP, int label_bytes) {
    return emtf::cvrp_lds_bytes(n, popP, label_bytes);
}

// label_bytes 2: split labels stored as uint16 (the host checks that every label of the
// batch, at most the cost of serving each customer alone, is below 2^16); else int32.
extern "C" int emtf_cvrp_launch(const emtf::CvrpArgs* args, const int* order, int n_inst, int threads, int lds_bytes,
                                int label_bytes, hipStream_t stream) {
    static bool lds_ok[2] = {false, false};
    const int w = label_bytes == 2 ? 0 : 1;
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
