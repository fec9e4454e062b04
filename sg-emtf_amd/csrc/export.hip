// export.hip -- elite export: the k best individuals of every task, ranked on the device.
//
// Replaces the reference's popSort + messageTransfer pair (proposed model/SG-EMTF.cu:462-486,
// :1601-1615), which bitonic-sorts 512 (fitness, pointer) pairs per task in a separate
// launch and copies the first BESTNUM_PERTASK rows back.  Here one workgroup per task ranks
// its population by counting in LDS -- order (fitness, index), NaN last, the order the
// engine uses everywhere -- and writes the first k rows to a dense [task][k] block.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "emtf_types.h"

namespace emtf {

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
        for (int e = tid; e < kk * D; e += blockDim.x) {
            const int r = e / D, d = e - r * D;
            out_x[((int64_t)t * k + r) * stride + d] = X[td.x_off + (int64_t)ord[r] * D + d];
        }
    }
    for (int r = tid; r < kk; r += blockDim.x) out_f[(int64_t)t * k + r] = fit[td.fit_off + ord[r]];
}

}  // namespace emtf

extern "C" int emtf_launch_elites(const emtf::TaskDev* tasks, int n_tasks, const float* X, const float* fit, int k,
                                  int stride, float* out_x, float* out_f, hipStream_t stream) {
    hipLaunchKernelGGL(emtf::elite_kernel, dim3(n_tasks), dim3(256), 1024 * 8, stream, tasks, X, fit, k, stride, out_x,
                       out_f);
    return (int)hipGetLastError();
}
