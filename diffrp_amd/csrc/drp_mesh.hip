// Vertex -> incident triangle-corner table (CSR), used by the deterministic
// vertex gathers of the triangle-centric backward passes.
//   inc_off [V+1]: exclusive prefix of corner counts per vertex
//   inc [3F]     : corner ids (3 t + k) grouped by vertex, ascending within a vertex
// Corners whose index is outside [0, V) are left out.  The order within a
// vertex comes from one radix sort of (vertex << 32 | corner) keys, so it is
// fixed whatever the vertex valence (a fan centre with 10^5 corners included).
#include <rocprim/device/device_radix_sort.hpp>

#include "drp_common.h"
#include "drp_kernels.h"

namespace drp {

__global__ void k_inc_zero(int *p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_inc_count(const int *__restrict__ tri, int F, int V, int *__restrict__ cnt) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 3ll * F) return;
    int v = tri[g];
    if ((unsigned)v < (unsigned)V) atomicAdd(&cnt[v], 1);
}

// exclusive scan, phase 1: per-block scan of 1024 elements, block totals out
__global__ void __launch_bounds__(1024) k_scan_block(const int *__restrict__ in, int *__restrict__ out, long long n,
                                                     int *__restrict__ block_sums) {
    __shared__ int wsum[16];
    long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    int v = i < n ? in[i] : 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int pre = 0;
    for (int w = 0; w < wave; ++w) pre += wsum[w];
    if (i < n) out[i] = pre + x - v;
    if (threadIdx.x == 1023) block_sums[blockIdx.x] = pre + x;
}

// phase 2: exclusive scan of the block totals by one block (sequential chunks)
__global__ void __launch_bounds__(1024) k_scan_sums(int *__restrict__ sums, int nb) {
    __shared__ int wsum[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int base = 0; base < nb; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < nb ? sums[i] : 0;
        int x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        int pre = 0, tot = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wave) pre += wsum[w];
            tot += wsum[w];
        }
        int c = carry;
        if (i < nb) sums[i] = c + pre + x - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
}

// phase 3: add block offsets; also writes the grand total at out[n]
__global__ void k_scan_add(int *__restrict__ out, long long n, const int *__restrict__ block_sums,
                           const int *__restrict__ in) {
    long long i = (long long)blockIdx.x * 1024 + threadIdx.x;
    if (i < n) {
        out[i] += block_sums[blockIdx.x];
        if (i == n - 1) out[n] = out[i] + in[i];
    }
}

__global__ void k_inc_keys(const int *__restrict__ tri, long long ne, int V, unsigned long long *__restrict__ keys) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ne) return;
    int v = tri[g];
    unsigned long long vv = (unsigned)v < (unsigned)V ? (unsigned)v : (unsigned)V;  // V: left out, sorted last
    keys[g] = (vv << 32) | (unsigned long long)g;
}

__global__ void k_inc_from_keys(const unsigned long long *__restrict__ keys, long long ne, int V,
                                int *__restrict__ inc) {
    long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ne) return;
    unsigned long long k = keys[g];
    if ((k >> 32) < (unsigned long long)V) inc[g] = (int)(k & 0xffffffffu);
}

static unsigned key_end_bit(int V) {
    unsigned b = 0;
    while ((1ull << b) <= (unsigned long long)V) ++b;  // bits of the sentinel V
    return 32 + b;
}

static size_t sort_temp_bytes(int V, long long ne) {
    size_t s = 0;
    if (ne <= 0) return 0;
    if (rocprim::radix_sort_keys(nullptr, s, (unsigned long long *)nullptr, (unsigned long long *)nullptr,
                                 (size_t)ne, 0, key_end_bit(V), (hipStream_t)0) != hipSuccess)
        return 0;
    return s;
}

static size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t incidence_ws_bytes(int V, int F) {
    long long nb = ((long long)V + 1023) / 1024 + 1;
    long long ne = 3ll * F;
    return a256((size_t)V * 4) + a256((size_t)nb * 4) + 2 * a256((size_t)ne * 8) + a256(sort_temp_bytes(V, ne));
}

int launch_incidence(const int *tri, int F, int V, int *inc_off, int *inc, void *ws, hipStream_t st) {
    if (V <= 0) return 0;
    char *b = (char *)ws;
    const long long ne = 3ll * F;
    int *cnt = (int *)b;
    int *bsum = (int *)(b + a256((size_t)V * 4));
    long long nb = ((long long)V + 1023) / 1024;
    char *kb = b + a256((size_t)V * 4) + a256((size_t)(nb + 1) * 4);
    unsigned long long *keys = (unsigned long long *)kb, *sorted = (unsigned long long *)(kb + a256((size_t)ne * 8));
    void *temp = kb + 2 * a256((size_t)ne * 8);
    hipLaunchKernelGGL(k_inc_zero, dim3((unsigned)(((long long)V + 255) / 256)), dim3(256), 0, st, cnt, (long long)V);
    if (ne) hipLaunchKernelGGL(k_inc_count, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, tri, F, V, cnt);
    hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(1024), 0, st, cnt, inc_off, (long long)V, bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, bsum, (int)nb);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(1024), 0, st, inc_off, (long long)V, bsum, cnt);
    if (ne) {
        hipLaunchKernelGGL(k_inc_keys, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, tri, ne, V, keys);
        size_t temp_bytes = sort_temp_bytes(V, ne);
        if (rocprim::radix_sort_keys(temp, temp_bytes, keys, sorted, (size_t)ne, 0, key_end_bit(V), st) !=
            hipSuccess)
            return -2;
        hipLaunchKernelGGL(k_inc_from_keys, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, sorted, ne, V, inc);
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace drp
