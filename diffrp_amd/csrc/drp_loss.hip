// Image L2 loss of the training step, sum over all elements of (a - b)^2, and
// its gradient 2 g (a - b): one read of both images forward, one read of both
// and one write of the gradient backward.  Or, when the forward already knows
// the gradient will be wanted, the forward also writes 2 (a - b) (the gradient
// for g = 1) in its pass and the backward scales it by g in place, a pass that
// reads g alone when g is 1 (the gradient of a loss that is itself the
// objective).
//
// The sum is deterministic: a grid fixed by the element count (at most
// kLossBlocks blocks), each thread summing its elements in double (two float4
// 256 apart per round, then grid-stride rounds), the block's threads reduced
// in a fixed tree, then one block adding the block sums in a fixed order.
#include "drp_common.h"
#include "drp_kernels.h"

namespace drp {

typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kLossBlocks = 65536;
constexpr int kLossPer = 2;  // float4 per thread and round
constexpr int kLossThreads = 256;

__device__ __forceinline__ double block_sum(double v, double *red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kLossThreads / 64; ++w) s += red[w];
    return s;
}

__global__ void __launch_bounds__(kLossThreads) k_l2_partial(const v4f *__restrict__ a,
                                                             const v4f *__restrict__ b, long long n4,
                                                             const float *__restrict__ a1,
                                                             const float *__restrict__ b1, long long n,
                                                             double *__restrict__ partial, v4f *__restrict__ ga,
                                                             float *__restrict__ ga1) {
    __shared__ double red[kLossThreads / 64];
    double s = 0.0;
    const long long round = (long long)gridDim.x * blockDim.x * kLossPer;
    for (long long i0 = (long long)blockIdx.x * blockDim.x * kLossPer + threadIdx.x; i0 < n4; i0 += round) {
        v4f x[kLossPer], y[kLossPer];
#pragma unroll
        for (int k = 0; k < kLossPer; ++k) {
            const long long i = i0 + (long long)k * blockDim.x;
            if (i < n4) {
                x[k] = __builtin_nontemporal_load(a + i);
                y[k] = __builtin_nontemporal_load(b + i);
            }
        }
#pragma unroll
        for (int k = 0; k < kLossPer; ++k) {
            const long long i = i0 + (long long)k * blockDim.x;
            if (i >= n4) break;
            const float d0 = x[k].x - y[k].x, d1 = x[k].y - y[k].y, d2 = x[k].z - y[k].z, d3 = x[k].w - y[k].w;
            s += (double)(d0 * d0) + (double)(d1 * d1) + (double)(d2 * d2) + (double)(d3 * d3);
            if (ga) {  // the gradient for g = 1, as k_l2_grad computes it: (2 g) (a - b)
                const v4f g = {2.0f * d0, 2.0f * d1, 2.0f * d2, 2.0f * d3};
                __builtin_nontemporal_store(g, ga + i);
            }
        }
    }
    // the tail that does not fill a float4
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = 4 * n4 + (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float d = a1[i] - b1[i];
        s += (double)(d * d);
        if (ga1) ga1[i] = 2.0f * d;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// the middle level of the sum: block b adds partials [b per, (b + 1) per) in a fixed order
constexpr int kLossMid = 256;  // blocks of the middle level
__global__ void __launch_bounds__(kLossThreads) k_l2_mid(const double *__restrict__ partial, int n_part, int per,
                                                         double *__restrict__ mid) {
    __shared__ double red[kLossThreads / 64];
    double s = 0.0;
    const int lo = blockIdx.x * per, hi = min(lo + per, n_part);
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) s += partial[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) mid[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kLossThreads) k_l2_finish(const double *__restrict__ partial, int n_part,
                                                            float *__restrict__ loss) {
    __shared__ double red[kLossThreads / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < n_part; i += blockDim.x) s += partial[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) *loss = (float)s;
}

__global__ void __launch_bounds__(kLossThreads) k_l2_grad(const v4f *__restrict__ a, const v4f *__restrict__ b,
                                                          long long n4, const float *__restrict__ a1,
                                                          const float *__restrict__ b1, long long n,
                                                          const float *__restrict__ grad_loss,
                                                          v4f *__restrict__ ga, v4f *__restrict__ gb,
                                                          float *__restrict__ ga1, float *__restrict__ gb1) {
    const float k = 2.0f * *grad_loss;  // d/da of g (a - b)^2
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) {
        const v4f x = __builtin_nontemporal_load(a + i), y = __builtin_nontemporal_load(b + i);
        const v4f g = {k * (x.x - y.x), k * (x.y - y.y), k * (x.z - y.z), k * (x.w - y.w)};
        if (ga) __builtin_nontemporal_store(g, ga + i);
        if (gb) __builtin_nontemporal_store(-g, gb + i);
    }
    const long long t = 4 * n4 + i;
    if (i < 4 && t < n) {
        const float g = k * (a1[t] - b1[t]);
        if (ga1) ga1[t] = g;
        if (gb1) gb1[t] = -g;
    }
}

__global__ void __launch_bounds__(kLossThreads) k_l2_grad_scalar(const float *__restrict__ a,
                                                                 const float *__restrict__ b, long long n,
                                                                 const float *__restrict__ grad_loss,
                                                                 float *__restrict__ ga, float *__restrict__ gb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = (2.0f * *grad_loss) * (a[i] - b[i]);
    if (ga) ga[i] = g;
    if (gb) gb[i] = -g;
}

// grad *= *grad_loss in place over a bounded grid (grid-stride); every block
// returns at once when it is 1
__global__ void __launch_bounds__(kLossThreads) k_l2_scale(v4f *__restrict__ g4, long long n4, float *__restrict__ g1,
                                                           long long n, const float *__restrict__ grad_loss) {
    const float k = *grad_loss;
    if (k == 1.0f) return;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const v4f x = __builtin_nontemporal_load(g4 + i);
        const v4f y = {x.x * k, x.y * k, x.z * k, x.w * k};
        __builtin_nontemporal_store(y, g4 + i);
    }
    for (long long t = 4 * n4 + (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) g1[t] *= k;
}

size_t l2_loss_ws_bytes() { return sizeof(double) * (kLossBlocks + kLossMid); }

int launch_l2_loss(const float *a, const float *b, long long n, float *loss, float *grad_a, void *ws,
                   hipStream_t st) {
    const long long n4 = ((((uintptr_t)a | (uintptr_t)b | (uintptr_t)grad_a) & 15) == 0) ? n / 4 : 0;
    long long want = (n4 + kLossThreads * kLossPer - 1) / (kLossThreads * kLossPer);
    const int blocks = (int)(want < 1 ? 1 : (want > kLossBlocks ? kLossBlocks : want));
    hipLaunchKernelGGL(k_l2_partial, dim3(blocks), dim3(kLossThreads), 0, st, (const v4f *)a, (const v4f *)b,
                       n4, a, b, n, (double *)ws, (v4f *)grad_a, grad_a);
    const double *part = (const double *)ws;
    int n_part = blocks;
    if (blocks > kLossThreads * 8) {  // many block sums: a middle level first
        const int per = (blocks + kLossMid - 1) / kLossMid;
        double *mid = (double *)ws + kLossBlocks;
        hipLaunchKernelGGL(k_l2_mid, dim3(kLossMid), dim3(kLossThreads), 0, st, part, blocks, per, mid);
        part = mid;
        n_part = kLossMid;
    }
    hipLaunchKernelGGL(k_l2_finish, dim3(1), dim3(kLossThreads), 0, st, part, n_part, loss);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_l2_loss_scale(float *grad, long long n, const float *grad_loss, hipStream_t st) {
    const long long n4 = (((uintptr_t)grad & 15) == 0) ? n / 4 : 0;  // unaligned: all of it in the scalar loop
    const long long items = n4 > 0 ? n4 : n;
    const long long want = (items + kLossThreads - 1) / kLossThreads;
    constexpr long long kScaleBlocks = 2048;  // few blocks: when the scalar is 1 they all return at once
    const int blocks = (int)(want < 1 ? 1 : (want > kScaleBlocks ? kScaleBlocks : want));
    hipLaunchKernelGGL(k_l2_scale, dim3(blocks), dim3(kLossThreads), 0, st, (v4f *)grad, n4, grad, n, grad_loss);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_l2_loss_bwd(const float *a, const float *b, long long n, const float *grad_loss, float *ga, float *gb,
                       hipStream_t st) {
    const bool vec = ((((uintptr_t)a | (uintptr_t)b | (uintptr_t)ga | (uintptr_t)gb) & 15) == 0);
    const long long n4 = vec ? n / 4 : 0;
    const long long items = n4 > 4 ? n4 : 4;  // the last 4 threads also take the tail
    const long long blocks = (items + kLossThreads - 1) / kLossThreads;
    if (!vec) {
        // scalar layout: every element is tail work of a one-element-per-thread grid
        const long long nb = (n + kLossThreads - 1) / kLossThreads;
        hipLaunchKernelGGL(k_l2_grad_scalar, dim3((unsigned)(nb < 1 ? 1 : nb)), dim3(kLossThreads), 0, st, a, b, n,
                           grad_loss, ga, gb);
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
    hipLaunchKernelGGL(k_l2_grad, dim3((unsigned)blocks), dim3(kLossThreads), 0, st, (const v4f *)a,
                       (const v4f *)b, n4, a, b, n, grad_loss, (v4f *)ga, (v4f *)gb, ga, gb);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
