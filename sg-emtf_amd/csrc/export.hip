// export.hip -- elite export (the k best individuals of every task, ranked on the device),
// the convergence trace (best / worst fitness of every task) and draw export (the engine's
// Philox blocks at given addresses).
//
// Replaces the reference's popSort + messageTransfer pair (proposed model/SG-EMTF.cu:462-486,
// :1601-1615), which bitonic-sorts 512 (fitness, pointer) pairs per task in a separate
// launch and copies the first BESTNUM_PERTASK rows back.  Here one workgroup per task ranks
// its population by counting in LDS -- order (fitness, index), NaN last, the order the
// engine uses everywhere -- and writes the first k rows to a dense [task][k] block.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "emtf_types.h"
#include "emtf_math.h"

namespace emtf {

// Staging of emtf_get_elites_noalloc: a fixed array in every GPU's memory for the life of
// the process.  Host code never uses its name as an address: emtf_scratch_resolve() asks
// the runtime for each device's copy once, with that device current, and callers pass the
// returned pointer to the kernels.
constexpr size_t kScratchFloats = size_t(1) << 20;  // 4 MiB
static __device__ __align__(256) float scratch[kScratchFloats];

__global__ __launch_bounds__(256) void elite_kernel(const TaskDev* __restrict__ tasks, const float* __restrict__ X,
                                                    const float* __restrict__ fit, int k, int stride,
                                                    float* __restrict__ out_x, float* __restrict__ out_f) {
    extern __shared__ float sfit[];
    int* ord = reinterpret_cast<int*>(sfit + 1024);
    const int t = blockIdx.x, tid = threadIdx.x;
    const TaskDev td = tasks[t];
    const int P = td.pop, D = td.dim;
    for (int i = tid; i < P; i += blockDim.x) {
        const float f = fit[td.fit_off + i];
        sfit[i] = (f != f) ? __builtin_inff() : f;
    }
    __syncthreads();
    for (int i = tid; i < P; i += blockDim.x) {
        const float fi = sfit[i];
        int r = 0;
        for (int p = 0; p < P; ++p) r += (sfit[p] < fi || (sfit[p] == fi && p < i)) ? 1 : 0;
        ord[r] = i;
    }
    __syncthreads();
    const int kk = k < P ? k : P;
    if (out_x) {
        // whole rows of the stride: genes, then zeros past the task's dimension
        for (int e = tid; e < kk * stride; e += blockDim.x) {
            const int r = e / stride, d = e - r * stride;
            out_x[((int64_t)t * k + r) * stride + d] = d < D ? X[td.x_off + (int64_t)ord[r] * D + d] : 0.0f;
        }
    }
    for (int r = tid; r < kk; r += blockDim.x) out_f[(int64_t)t * k + r] = fit[td.fit_off + ord[r]];
}

// One Philox4x32-10 block per address {task, generation, purpose, index, sub, block},
// through the engine's own Rng (the call the epoch kernels make).
__global__ __launch_bounds__(256) void draws_kernel(uint32_t seed_lo, uint32_t seed_hi, int n,
                                                    const uint32_t* __restrict__ addr, uint32_t* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t* a = addr + 6 * (int64_t)i;
    const Rng rng{seed_lo, seed_hi, a[0]};
    const U4 b = rng.block(a[1], a[2], a[3], a[4], a[5]);
    uint4 v;
    v.x = b.x;
    v.y = b.y;
    v.z = b.z;
    v.w = b.w;
    reinterpret_cast<uint4*>(out)[i] = v;
}

// Best (and, with out_worst, worst) fitness of every task, NaN ignored (NaN when every
// individual is NaN); one workgroup per task.  min / max are exact in any order.
__global__ __launch_bounds__(256) void trace_kernel(const TaskDev* __restrict__ tasks, const float* __restrict__ fit,
                                                    float* __restrict__ out_best, float* __restrict__ out_worst) {
    __shared__ float smin[256], smax[256];
    const int t = blockIdx.x, tid = threadIdx.x;
    const TaskDev td = tasks[t];
    const float nan = __builtin_nanf("");
    float mn = nan, mx = nan;
    for (int i = tid; i < td.pop; i += 256) {
        const float f = fit[td.fit_off + i];
        mn = __builtin_fminf(mn, f);
        mx = __builtin_fmaxf(mx, f);
    }
    smin[tid] = mn;
    smax[tid] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            smin[tid] = __builtin_fminf(smin[tid], smin[tid + w]);
            smax[tid] = __builtin_fmaxf(smax[tid], smax[tid + w]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        out_best[t] = smin[0];
        if (out_worst) out_worst[t] = smax[0];
    }
}

// The transfer donors of online admission at the boundary g: the ids, ascending, of the
// tasks that ran up to g and run on after it (gen_base < g < gen_end), then their count
// at list[n_tasks].  One workgroup of 1024 threads, 1024 tasks a round: each wave counts
// its members with a ballot, the waves' counts give each member its slot.
__global__ __launch_bounds__(1024) void donor_kernel(const TaskDev* __restrict__ tasks, int n_tasks, int g,
                                                     int32_t* __restrict__ list) {
    __shared__ int wsum[16];
    __shared__ int done;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) done = 0;
    __syncthreads();
    for (int c0 = 0; c0 < n_tasks; c0 += 1024) {
        const int t = c0 + tid;
        bool in = false;
        if (t < n_tasks) {
            const TaskDev td = tasks[t];
            in = td.gen_base < g && g < td.gen_end;
        }
        const unsigned long long bal = __ballot(in);
        if (lane == 0) wsum[w] = __popcll(bal);
        __syncthreads();
        int off = done;
        for (int q = 0; q < w; ++q) off += wsum[q];
        if (in) list[off + __popcll(bal & ((1ull << lane) - 1ull))] = t;
        __syncthreads();
        if (tid == 0) {
            int tot = 0;
            for (int q = 0; q < 16; ++q) tot += wsum[q];
            done += tot;
        }
        __syncthreads();
    }
    if (tid == 0) list[n_tasks] = done;
}

// The device generation counter of replayed epochs: set (after directly launched
// epochs) or advanced by one epoch (the last node of a captured epoch).
__global__ void gen_counter_kernel(int32_t* gen, int32_t value, int32_t add) {
    if (threadIdx.x == 0) gen[0] = add ? gen[0] + value : value;
}

// n int32 values set to v (the persistent schedule's counters; a kernel, not a memset, so
// the call can sit in a captured graph).
__global__ void fill_i32_kernel(int32_t* p, int n, int32_t v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace emtf

extern "C" int emtf_launch_fill_i32(int32_t* p, int n, int32_t value, hipStream_t stream) {
    if (n <= 0) return 0;
    const int blocks = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
    hipLaunchKernelGGL(emtf::fill_i32_kernel, dim3(blocks), dim3(256), 0, stream, p, n, value);
    return (int)hipGetLastError();
}

extern "C" int emtf_launch_gen_counter(int32_t* gen, int32_t value, int32_t add, hipStream_t stream) {
    hipLaunchKernelGGL(emtf::gen_counter_kernel, dim3(1), dim3(64), 0, stream, gen, value, add);
    return (int)hipGetLastError();
}

extern "C" int emtf_launch_donors(const emtf::TaskDev* tasks, int n_tasks, int g, int32_t* list,
                                  hipStream_t stream) {
    hipLaunchKernelGGL(emtf::donor_kernel, dim3(1), dim3(1024), 0, stream, tasks, n_tasks, g, list);
    return (int)hipGetLastError();
}

extern "C" int emtf_launch_trace(const emtf::TaskDev* tasks, int n_tasks, const float* fit, float* out_best,
                                 float* out_worst, hipStream_t stream) {
    hipLaunchKernelGGL(emtf::trace_kernel, dim3(n_tasks), dim3(256), 0, stream, tasks, fit, out_best, out_worst);
    return (int)hipGetLastError();
}

extern "C" int emtf_launch_draws(uint64_t seed, int n, const uint32_t* addr, uint32_t* out, hipStream_t stream) {
    hipLaunchKernelGGL(emtf::draws_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, (uint32_t)seed,
                       (uint32_t)(seed >> 32), n, addr, out);
    return (int)hipGetLastError();
}

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
float* g_scratch[kMaxDevices];  // each device's copy of emtf::scratch, once resolved
}  // namespace

// Device address of `device`'s staging array (resolved once per device, with that device
// current) and its capacity in floats.  Returns a hipError_t value.
extern "C" int emtf_scratch_resolve(int device, float** out, size_t* capacity) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
    if (!g_scratch[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != device) e = hipSetDevice(device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(emtf::scratch));
        if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return (int)e;
        g_scratch[device] = static_cast<float*>(p);
    }
    *out = g_scratch[device];
    *capacity = emtf::kScratchFloats;
    return (int)hipSuccess;
}

extern "C" int emtf_launch_elites(const emtf::TaskDev* tasks, int n_tasks, const float* X, const float* fit, int k,
                                  int stride, float* out_x, float* out_f, hipStream_t stream) {
    hipLaunchKernelGGL(emtf::elite_kernel, dim3(n_tasks), dim3(256), 1024 * 8, stream, tasks, X, fit, k, stride, out_x,
                       out_f);
    return (int)hipGetLastError();
}
