// nf_scan.hip - exclusive prefix sum over u32 arrays and fill kernels.
//
// Replaces the thrust::inclusive_scan / exclusive_scan calls of the
// reference (src/get_contact_arrays.cu:101, src/conductivity.cu:495,
// src/adjacency_graph.cu:79) and its in-block work-efficient scan
// (inc/sort.h:103-154).  Three passes (tile sums, scan of tile sums, tile
// rescan) with 2048-element tiles, 256 threads x 8 items, or for arrays of
// at most 16384 one pass of one workgroup; fixed launch shape for a given
// n, so the passes can be captured in a HIP graph.
#include "nf_internal.h"

namespace nf {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, WAVE);
    v += (lane >= o) ? t : 0u;
  }
  return v;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sm, uint32_t* total) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == WAVE - 1) sm[wv] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / WAVE; ++w) {
    const uint32_t s = sm[w];
    if (w < wv) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(const uint32_t* __restrict__ in,
                                                             uint64_t n, uint32_t* __restrict__ sums) {
  __shared__ uint32_t sm[SCAN_THREADS / WAVE];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint64_t i = base + k;
    const uint32_t v = in[i < n ? i : n - 1];  // clamped load, masked below
    s += i < n ? v : 0u;
  }
  uint32_t tot;
  block_excl_scan(s, sm, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_sums(uint32_t* __restrict__ sums,
                                                            uint32_t nt, uint32_t* __restrict__ out,
                                                            uint64_t n) {
  __shared__ uint32_t sm[SCAN_THREADS / WAVE];
  __shared__ uint32_t dummy[SCAN_THREADS];
  uint32_t carry = 0;
  for (uint32_t b = 0; b < nt; b += SCAN_THREADS) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t v = sums[i < nt ? i : nt - 1];
    uint32_t tot;
    const uint32_t ex = block_excl_scan(i < nt ? v : 0u, sm, &tot);
    // branch-free store: out-of-range threads write a private LDS slot
    uint32_t* dst = i < nt ? &sums[i] : &dummy[threadIdx.x];
    *dst = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) out[n] = carry;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_apply(const uint32_t* __restrict__ in,
                                                             uint64_t n,
                                                             const uint32_t* __restrict__ sums,
                                                             uint32_t* __restrict__ out) {
  __shared__ uint32_t sm[SCAN_THREADS / WAVE];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint64_t i = base + k;
    const uint32_t x = in[i < n ? i : n - 1];
    v[k] = i < n ? x : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(s, sm, &tot) + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

// Short arrays (the cells of a single packing): one workgroup of 1024
// threads walks the array in tiles of 4096 with a running carry, one launch
// instead of three.
constexpr int SCAN1_THREADS = 1024;
constexpr int SCAN1_ITEMS = 4;
constexpr uint64_t SCAN1_MAX = 16384;
__global__ void __launch_bounds__(SCAN1_THREADS) k_scan_one(const uint32_t* __restrict__ in,
                                                            uint64_t n, uint32_t* __restrict__ out) {
  __shared__ uint32_t sm[SCAN1_THREADS / WAVE];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = threadIdx.x / WAVE;
  uint32_t carry = 0;
  for (uint64_t t0 = 0; t0 < n; t0 += (uint64_t)SCAN1_THREADS * SCAN1_ITEMS) {
    const uint64_t base = t0 + (uint64_t)threadIdx.x * SCAN1_ITEMS;
    uint32_t v[SCAN1_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN1_ITEMS; ++k) {
      const uint64_t i = base + k;
      const uint32_t x = in[i < n ? i : n - 1];  // clamped load, masked below
      v[k] = i < n ? x : 0u;
      sum += v[k];
    }
    const uint32_t inc = wave_incl_scan(sum);
    if (lane == WAVE - 1) sm[wv] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN1_THREADS / WAVE; ++w) {
      const uint32_t x = sm[w];
      off += w < wv ? x : 0u;
      tot += x;
    }
    __syncthreads();
    uint32_t run = carry + off + inc - sum;
#pragma unroll
    for (int k = 0; k < SCAN1_ITEMS; ++k) {
      const uint64_t i = base + k;
      if (i < n) out[i] = run;
      run += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) out[n] = carry;
}

size_t scan_tmp_words(uint64_t n) { return (size_t)((n + SCAN_TILE - 1) / SCAN_TILE) + 1; }

// out[0..n) = exclusive prefix of in, out[n] = total.  in and out may alias.
void scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s) {
  if (n > 0 && n <= SCAN1_MAX) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(SCAN1_THREADS), 0, s, in, n, out);
    return;
  }
  const uint32_t nt = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
  if (nt == 0) {
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, s, tmp, 0u, out, (uint64_t)0);
    return;
  }
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(SCAN_THREADS), 0, s, in, n, tmp);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, s, tmp, nt, out, n);
  hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(SCAN_THREADS), 0, s, in, n, tmp, out);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, s, p, v, n);
}

}  // namespace nf
