// Stencil partition of a layer's pixels by owning object, for the
// stencil-masked G-buffer path (diffrp/rendering/surface_deferred.py:283-290,
// interpolator.py:87-97).  diffrp takes owner = stencils[tri_id] per pixel and
// one mask.nonzero() per object, i.e. one host synchronisation per object and
// layer.  Here one radix sort of (owner << 32 | pixel) keys orders every
// pixel by owner and, within an owner, by pixel -- the order nonzero() gives --
// and the boundaries of the sorted runs give each owner's [start, end); the
// caller reads starts/ends once per layer.
#include <rocprim/device/device_radix_sort.hpp>

#include "drp_common.h"
#include "drp_kernels.h"

namespace drp {

__global__ void k_stencil_keys(const float4 *__restrict__ rast, long long npx, const int *__restrict__ stencils,
                               int F, int n_owner, unsigned long long *__restrict__ keys) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx) return;
    float w = rast[p].w;
    int id = w > 0.f ? float_to_triidx(w) : 0;  // triangle + 1, 0: background
    int o = (id >= 0 && id <= F) ? stencils[id] : 0;
    if (o < 0 || o > n_owner) o = 0;
    keys[p] = ((unsigned long long)(unsigned)o << 32) | (unsigned long long)p;
}

__global__ void k_stencil_runs(const unsigned long long *__restrict__ sorted, long long npx, int *__restrict__ order,
                               int *__restrict__ starts, int *__restrict__ ends) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npx) return;
    unsigned long long k = sorted[i];
    int o = (int)(k >> 32);
    order[i] = (int)(k & 0xffffffffu);
    if (i == 0 || (int)(sorted[i - 1] >> 32) != o) starts[o] = (int)i;
    if (i == npx - 1 || (int)(sorted[i + 1] >> 32) != o) ends[o] = (int)(i + 1);
}

__global__ void k_zero2_i32(int *a, int *b, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = b[i] = 0;
}

static unsigned owner_end_bit(int n_owner) {
    unsigned b = 0;
    while ((1ull << b) <= (unsigned long long)n_owner) ++b;
    return 32 + (b ? b : 1);
}

static size_t sort_bytes(long long npx, int n_owner) {
    size_t s = 0;
    if (npx <= 0) return 0;
    if (rocprim::radix_sort_keys(nullptr, s, (unsigned long long *)nullptr, (unsigned long long *)nullptr,
                                 (size_t)npx, 0, owner_end_bit(n_owner), (hipStream_t)0) != hipSuccess)
        return 0;
    return s;
}

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t stencil_partition_ws_bytes(long long npx, int n_owner) {
    return 2 * a256((size_t)npx * 8) + a256(sort_bytes(npx, n_owner));
}

int launch_stencil_partition(const float *rast, long long npx, const int *stencils, int F, int n_owner, int *order,
                             int *starts, int *ends, void *ws, hipStream_t st) {
    char *b = (char *)ws;
    unsigned long long *keys = (unsigned long long *)b, *sorted = (unsigned long long *)(b + a256((size_t)npx * 8));
    void *temp = b + 2 * a256((size_t)npx * 8);
    hipLaunchKernelGGL(k_zero2_i32, dim3((unsigned)((n_owner + 1 + 255) / 256)), dim3(256), 0, st, starts, ends,
                       n_owner + 1);
    if (npx == 0) return hipGetLastError() == hipSuccess ? 0 : -2;
    dim3 grid((unsigned)((npx + 255) / 256));
    hipLaunchKernelGGL(k_stencil_keys, grid, dim3(256), 0, st, reinterpret_cast<const float4 *>(rast), npx, stencils,
                       F, n_owner, keys);
    size_t temp_bytes = sort_bytes(npx, n_owner);
    if (rocprim::radix_sort_keys(temp, temp_bytes, keys, sorted, (size_t)npx, 0, owner_end_bit(n_owner), st) !=
        hipSuccess)
        return -2;
    hipLaunchKernelGGL(k_stencil_runs, grid, dim3(256), 0, st, sorted, npx, order, starts, ends);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
