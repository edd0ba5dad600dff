// nf_post.hip - cell-list build for a given packing, canonical ordering of
// contact rows, and exact int64 histograms.
#include "nf_internal.h"

namespace nf {

static int grid_for(uint64_t n, int threads) {
  uint64_t b = (n + threads - 1) / threads;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

// ---------------------------------------------------------------------------
// Cell list of a packing (the role of writeGlbMem's per-cell placement,
// inc/memoryIO.h:361-429, for an existing packing) as a two-level counting
// sort with no device-scope atomics: those resolve beyond the XCD's L2 and
// cap a one-atomic-per-rod binning at a few tens of G/s.  Buckets are runs
// of 2^shift consecutive cells in linear cell order, so a bucket's cells
// own one contiguous range of the output.
//   1. k_bin_count: each block counts its chunk of rods per bucket in LDS
//      and writes its column of the bucket-major matrix M[bucket][block];
//   2. scan_u32 of M gives each (bucket, block) its range in bucket order;
//   3. k_bin_scatter: each block places its rods there (LDS cursors);
//   4. k_bin_cells: one workgroup per bucket counts its rods per cell in
//      LDS, scans the counts, writes the cells' start, count and largest
//      half length, and places the rods in cell order.
// The order of the rods inside a cell is not fixed (LDS atomics); every
// pass reading the list is independent of it.
// ---------------------------------------------------------------------------
constexpr int BIN_THREADS = 1024;

struct BinArg {
  int32_t nc[3];
  double sc[3];
  uint32_t shift;  // log2 of the cells per bucket
  uint32_t nb;     // buckets
  uint32_t nblk;   // blocks of passes 1 and 3
  uint32_t chunk;  // rods per block of passes 1 and 3
  uint32_t ncells;
};

__global__ void __launch_bounds__(BIN_THREADS) k_bin_count(const float4* __restrict__ P, uint32_t n, BinArg g,
                                                           uint32_t* __restrict__ M) {
  extern __shared__ uint32_t hist[];
  for (uint32_t b = threadIdx.x; b < g.nb; b += BIN_THREADS) hist[b] = 0u;
  __syncthreads();
  const uint64_t lo = (uint64_t)blockIdx.x * g.chunk;
  const uint64_t hi = lo + g.chunk < (uint64_t)n ? lo + g.chunk : (uint64_t)n;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += BIN_THREADS)
    atomicAdd(&hist[cell_index(P[i], g.nc, g.sc) >> g.shift], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < g.nb; b += BIN_THREADS)
    M[(uint64_t)b * g.nblk + blockIdx.x] = hist[b];
}

__global__ void __launch_bounds__(BIN_THREADS) k_bin_scatter(const float4* __restrict__ P,
                                                             const float4* __restrict__ D, uint32_t n,
                                                             BinArg g, const uint32_t* __restrict__ M,
                                                             float4* __restrict__ tP, float4* __restrict__ tD,
                                                             uint32_t* __restrict__ tid) {
  extern __shared__ uint32_t cur[];
  for (uint32_t b = threadIdx.x; b < g.nb; b += BIN_THREADS) cur[b] = M[(uint64_t)b * g.nblk + blockIdx.x];
  __syncthreads();
  const uint64_t lo = (uint64_t)blockIdx.x * g.chunk;
  const uint64_t hi = lo + g.chunk < (uint64_t)n ? lo + g.chunk : (uint64_t)n;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += BIN_THREADS) {
    const float4 p = P[i];
    const uint32_t dst = atomicAdd(&cur[cell_index(p, g.nc, g.sc) >> g.shift], 1u);
    tP[dst] = p;
    tD[dst] = D[i];
    tid[dst] = (uint32_t)i;
  }
}

__global__ void __launch_bounds__(BIN_THREADS) k_bin_cells(const float4* __restrict__ tP,
                                                           const float4* __restrict__ tD,
                                                           const uint32_t* __restrict__ tid, BinArg g,
                                                           const uint32_t* __restrict__ M,
                                                           float4* __restrict__ oP, float4* __restrict__ oD,
                                                           uint32_t* __restrict__ oid,
                                                           uint32_t* __restrict__ start,
                                                           uint32_t* __restrict__ count,
                                                           float* __restrict__ hl) {
  extern __shared__ uint32_t lds[];
  const uint32_t K = 1u << g.shift;
  uint32_t* cnt = lds;      // rods per cell
  uint32_t* mx = lds + K;   // largest half length per cell (float bits, >= 0)
  uint32_t* off = mx + K;   // exclusive prefix, then the placement cursor
  __shared__ uint32_t wsum[BIN_THREADS / WAVE];
  const uint32_t b = blockIdx.x;
  const uint32_t base = M[(uint64_t)b * g.nblk];
  const uint32_t end = M[(uint64_t)(b + 1) * g.nblk];
  const uint32_t c0 = b << g.shift;
  for (uint32_t k = threadIdx.x; k < K; k += BIN_THREADS) {
    cnt[k] = 0u;
    mx[k] = 0u;
  }
  __syncthreads();
  for (uint32_t i = base + threadIdx.x; i < end; i += BIN_THREADS) {
    const float4 p = tP[i];
    const uint32_t k = cell_index(p, g.nc, g.sc) - c0;
    atomicAdd(&cnt[k], 1u);
    atomicMax(&mx[k], __float_as_uint(p.w));
  }
  __syncthreads();
  // exclusive scan of cnt: each thread sums a run of K / BIN_THREADS cells
  // (at least one), then a wave scan and the wave totals
  const uint32_t per = (K + BIN_THREADS - 1) / BIN_THREADS;
  const uint32_t k0 = threadIdx.x * per;
  uint32_t sum = 0;
  for (uint32_t k = k0; k < k0 + per && k < K; ++k) sum += cnt[k];
  const int lane = threadIdx.x & (WAVE - 1);
  uint32_t inc = sum;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, WAVE);
    inc += lane >= o ? t : 0u;
  }
  if (lane == WAVE - 1) wsum[threadIdx.x / WAVE] = inc;
  __syncthreads();
  uint32_t run = inc - sum;
  for (int w = 0; w < (int)(threadIdx.x / WAVE); ++w) run += wsum[w];
  for (uint32_t k = k0; k < k0 + per && k < K; ++k) {
    off[k] = run;
    run += cnt[k];
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < K; k += BIN_THREADS) {
    const uint32_t c = c0 + k;
    if (c < g.ncells) {
      start[c] = base + off[k];
      count[c] = cnt[k];
      hl[c] = __uint_as_float(mx[k]);
    }
  }
  if (b + 1 == g.nb && threadIdx.x == 0) start[g.ncells] = end;
  __syncthreads();
  for (uint32_t i = base + threadIdx.x; i < end; i += BIN_THREADS) {
    const float4 p = tP[i];
    const uint32_t dst = base + atomicAdd(&off[cell_index(p, g.nc, g.sc) - c0], 1u);
    oP[dst] = p;
    oD[dst] = tD[i];
    oid[dst] = tid[i];
  }
}

// Buckets of 2^shift cells, shift the smallest with at most BIN_BUCKETS
// buckets (at least 32 cells, at most 4096: 48 KB of LDS per bucket
// workgroup), and about BIN_BLOCKS blocks of at least 2048 rods for the
// histogram and scatter passes, so that each (bucket, block) range written
// by the scatter holds a few tens of rods.
constexpr uint32_t BIN_BUCKETS = 1024;
constexpr uint32_t BIN_BLOCKS = 256;

BinPlan bin_plan(uint32_t n, uint64_t ncells) {
  BinPlan bp{};
  uint32_t shift = 5;
  while (shift < 12 && ((ncells + (1ull << shift) - 1) >> shift) > BIN_BUCKETS) ++shift;
  bp.shift = shift;
  bp.nb = (uint32_t)((ncells + (1ull << shift) - 1) >> shift);
  uint64_t chunk = ((uint64_t)n + BIN_BLOCKS - 1) / BIN_BLOCKS;
  if (chunk < 2048) chunk = 2048;
  chunk = (chunk + BIN_THREADS - 1) / BIN_THREADS * BIN_THREADS;
  bp.chunk = (uint32_t)chunk;
  bp.nblk = n == 0 ? 0u : (uint32_t)(((uint64_t)n + chunk - 1) / chunk);
  bp.m_words = (uint64_t)bp.nb * bp.nblk;
  return bp;
}

void launch_bin_rods(const BinPlan& bp, const float4* P, const float4* D, uint32_t n, const int32_t ncell[3],
                     const double scale[3], uint32_t* M, uint32_t* scan_tmp, float4* tP, float4* tD,
                     uint32_t* tid, float4* oP, float4* oD, uint32_t* oid, uint32_t* start, uint32_t* count,
                     float* hl, hipStream_t s) {
  BinArg g;
  for (int a = 0; a < 3; ++a) {
    g.nc[a] = ncell[a];
    g.sc[a] = scale[a];
  }
  g.shift = bp.shift;
  g.nb = bp.nb;
  g.nblk = bp.nblk;
  g.chunk = bp.chunk;
  g.ncells = (uint32_t)((uint64_t)ncell[0] * ncell[1] * ncell[2]);
  if (n == 0) {
    fill_u32(start, 0u, (uint64_t)g.ncells + 1, s);
    fill_u32(count, 0u, g.ncells, s);
    fill_u32((uint32_t*)hl, 0u, g.ncells, s);
    return;
  }
  const size_t hist_bytes = sizeof(uint32_t) * bp.nb;
  hipLaunchKernelGGL(k_bin_count, dim3(bp.nblk), dim3(BIN_THREADS), hist_bytes, s, P, n, g, M);
  scan_u32(M, M, bp.m_words, scan_tmp, s);
  hipLaunchKernelGGL(k_bin_scatter, dim3(bp.nblk), dim3(BIN_THREADS), hist_bytes, s, P, D, n, g, M, tP, tD,
                     tid);
  hipLaunchKernelGGL(k_bin_cells, dim3(bp.nb), dim3(BIN_THREADS), sizeof(uint32_t) * 3u * (1u << bp.shift), s,
                     tP, tD, tid, g, M, oP, oD, oid, start, count, hl);
}

// ---------------------------------------------------------------------------
// Contact rows: sort each row by (J, image) so the list is canonical whatever
// order the record pass found the pairs in.  Rows are short (a few entries),
// so one thread per row with an insertion sort in global memory.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool img_less(int32_t a, int32_t b) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int x = (int)(int8_t)((a >> (8 * d)) & 0xFF);
    const int y = (int)(int8_t)((b >> (8 * d)) & 0xFF);
    if (x != y) return x < y;
  }
  return false;
}

__global__ void k_row_sort(const uint32_t* __restrict__ rowptr, uint32_t n, uint32_t* __restrict__ J,
                           float* __restrict__ gap, double* __restrict__ gap64,
                           int32_t* __restrict__ img) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    const uint32_t b = rowptr[r], e = rowptr[r + 1];
    for (uint32_t i = b + 1; i < e; ++i) {
      const uint32_t kj = J[i];
      const int32_t ki = img[i];
      const float kg = gap[i];
      const double kd = gap64[i];
      uint32_t k = i;
      while (k > b && (J[k - 1] > kj || (J[k - 1] == kj && img_less(ki, img[k - 1])))) {
        J[k] = J[k - 1];
        img[k] = img[k - 1];
        gap[k] = gap[k - 1];
        gap64[k] = gap64[k - 1];
        --k;
      }
      J[k] = kj;
      img[k] = ki;
      gap[k] = kg;
      gap64[k] = kd;
    }
  }
}

void launch_row_sort(const uint32_t* rowptr, uint32_t n, uint32_t* J, float* gap, double* gap64,
                     int32_t* img, hipStream_t s) {
  hipLaunchKernelGGL(k_row_sort, dim3(grid_for(n, 256)), dim3(256), 0, s, rowptr, n, J, gap,
                     gap64, img);
}

// Contact area of every contact, one thread per row (rows are short).
__global__ void k_contact_area(const float4* __restrict__ P, const float4* __restrict__ D, uint32_t n,
                               const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ J,
                               const int32_t* __restrict__ img, double bx, double by, double bz,
                               double rlim, double* __restrict__ area) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t b = rowptr[i], e = rowptr[i + 1];
    if (b == e) continue;
    const float4 P1 = P[i], D1 = D[i];
    for (uint32_t k = b; k < e; ++k) {
      const uint32_t j = J[k];
      const int32_t m = img[k];
      const double sx = (double)(int)(int8_t)(m & 0xFF) * bx;
      const double sy = (double)(int)(int8_t)((m >> 8) & 0xFF) * by;
      const double sz = (double)(int)(int8_t)((m >> 16) & 0xFF) * bz;
      area[k] = contact_area_f64(P1, D1, P[j], D[j], sx, sy, sz, rlim);
    }
  }
}

void launch_contact_area(const float4* P, const float4* D, uint32_t n, const uint32_t* rowptr,
                         const uint32_t* J, const int32_t* img, const double box[3], double rlim,
                         double* area, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_contact_area, dim3(grid_for(n, 256)), dim3(256), 0, s, P, D, n, rowptr, J, img,
                     box[0], box[1], box[2], rlim, area);
}

// CSR over I from the unsorted contact list (the pair pass appends in any
// order; the row sort makes every row canonical afterwards).
__global__ void k_csr_count(const uint32_t* __restrict__ I, uint64_t c, uint32_t* __restrict__ cnt) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < c;
       k += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[I[k]], 1u);
}

__global__ void k_csr_scatter(const uint32_t* __restrict__ I, const uint32_t* __restrict__ J,
                              const float* __restrict__ gap, const double* __restrict__ gap64,
                              const int32_t* __restrict__ img, uint64_t c,
                              const uint32_t* __restrict__ rowptr, uint32_t* __restrict__ fill,
                              uint32_t* __restrict__ oJ, float* __restrict__ ogap,
                              double* __restrict__ ogap64, int32_t* __restrict__ oimg) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < c;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = I[k];
    const uint32_t pos = rowptr[i] + atomicAdd(&fill[i], 1u);
    oJ[pos] = J[k];
    ogap[pos] = gap[k];
    ogap64[pos] = gap64[k];
    oimg[pos] = img[k];
  }
}

void launch_csr_count(const uint32_t* I, uint64_t c, uint32_t* cnt, hipStream_t s) {
  if (c == 0) return;
  hipLaunchKernelGGL(k_csr_count, dim3(grid_for(c, 256)), dim3(256), 0, s, I, c, cnt);
}

void launch_csr_scatter(const uint32_t* I, const uint32_t* J, const float* gap, const double* gap64,
                        const int32_t* img, uint64_t c, const uint32_t* rowptr, uint32_t* fill,
                        uint32_t* oJ, float* ogap, double* ogap64, int32_t* oimg, hipStream_t s) {
  if (c == 0) return;
  hipLaunchKernelGGL(k_csr_scatter, dim3(grid_for(c, 256)), dim3(256), 0, s, I, J, gap, gap64, img, c,
                     rowptr, fill, oJ, ogap, ogap64, oimg);
}

// ---------------------------------------------------------------------------
// Centre-centre distance statistics (statKernel_CENT, inc/statKernel_CENT.h:
// 8-122, driven by collectStats, src/statistics_collection.cu:866-1033):
// every unordered pair of distinct rods over all periodic images with centre
// distance below bins * width, binned by comparing d^2 with the squared bin
// edges, plus exact fixed-point sums of round(d 2^24) and round(d^2 2^16).
// One workgroup (one wave) per home cell over a half stencil; each neighbour
// cell streams through LDS in 64-rod tiles; the FP64 distance is evaluated
// with the lower id first and the image applied to the higher id, so every
// bin decision matches a sequential evaluation.
// ---------------------------------------------------------------------------
constexpr int CENT_MAX_BINS = 1024;

// wave-uniform copies (SGPR) of values every lane already agrees on, so the
// loops that hold barriers have scalar trip counts
__device__ __forceinline__ uint32_t uni_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

struct CentArg {
  const float4* P;       // compact cell list (positions, half length)
  const uint32_t* X;     // rod ids
  const uint32_t* start;
  const uint32_t* count;
  const int4* st;
  uint32_t st_n;
  int32_t nc[3];
  double box[3];
  const double* e2;      // bins + 1 squared edges
  int bins;
  float r2f;             // FP32 prefilter bound
  unsigned long long* out;  // bins counts, then pairs, sum_q, sum2_q
};

__global__ void __launch_bounds__(WAVE) k_cent(CentArg a) {
  __shared__ float4 tP[WAVE];
  __shared__ uint32_t tX[WAVE];
  __shared__ uint32_t hist[CENT_MAX_BINS];
  const int lane = threadIdx.x;
  for (int k = lane; k < a.bins; k += WAVE) hist[k] = 0u;
  __syncthreads();
  unsigned long long cnt = 0, sq = 0, s2q = 0;
  const uint32_t ncells = (uint32_t)a.nc[0] * (uint32_t)a.nc[1] * (uint32_t)a.nc[2];
  const double e2max = a.e2[a.bins];
  for (uint32_t cell = blockIdx.x; cell < ncells; cell += gridDim.x) {
    const uint32_t pc = uni_u32(a.count[cell]), ps = uni_u32(a.start[cell]);
    const int c0 = (int)(cell % (uint32_t)a.nc[0]);
    const int c1 = (int)((cell / (uint32_t)a.nc[0]) % (uint32_t)a.nc[1]);
    const int c2 = (int)(cell / ((uint32_t)a.nc[0] * (uint32_t)a.nc[1]));
    for (uint32_t pb = 0; pb < pc; pb += WAVE) {
      const bool valid = pb + (uint32_t)lane < pc;
      const uint32_t pi = ps + min(pb + (uint32_t)lane, pc - 1u);
      const float4 P1 = a.P[pi];
      const uint32_t X1 = a.X[pi];
      for (uint32_t si = 0; si < a.st_n; ++si) {
        const int4 so = a.st[si];
        const int raw[3] = {c0 + so.x, c1 + so.y, c2 + so.z};
        int img[3], m[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          int mm = raw[d] % a.nc[d];
          if (mm < 0) mm += a.nc[d];
          img[d] = (raw[d] - mm) / a.nc[d];
          m[d] = mm;
        }
        const uint32_t nc = (uint32_t)(m[0] + a.nc[0] * (m[1] + a.nc[1] * m[2]));
        const uint32_t n = uni_u32(a.count[nc]), ns = uni_u32(a.start[nc]);
        const bool same = so.x == 0 && so.y == 0 && so.z == 0;
        const double sx = (double)img[0] * a.box[0], sy = (double)img[1] * a.box[1],
                     sz = (double)img[2] * a.box[2];
        const float fx = (float)sx, fy = (float)sy, fz = (float)sz;
        for (uint32_t tb = 0; tb < n; tb += WAVE) {
          const uint32_t tn = min((uint32_t)WAVE, n - tb);
          const uint32_t q = ns + tb + min((uint32_t)lane, tn - 1u);
          tP[lane] = a.P[q];
          tX[lane] = a.X[q];
          __syncthreads();
          if (valid) {
            for (uint32_t j = 0; j < tn; ++j) {
              if (same && tb + j <= pb + (uint32_t)lane) continue;
              const uint32_t xj = tX[j];
              if (xj == X1) continue;  // a rod and its own image
              const float4 P2 = tP[j];
              const float wx = (P1.x - P2.x) - fx, wy = (P1.y - P2.y) - fy, wz = (P1.z - P2.z) - fz;
              if (fmaf(wz, wz, fmaf(wy, wy, wx * wx)) > a.r2f) continue;
              const bool swap = X1 > xj;
              const float4 Pa = swap ? P2 : P1, Pb = swap ? P1 : P2;
              const double ux = swap ? -sx : sx, uy = swap ? -sy : sy, uz = swap ? -sz : sz;
              const double dx = ((double)Pa.x - (double)Pb.x) - ux;
              const double dy = ((double)Pa.y - (double)Pb.y) - uy;
              const double dz = ((double)Pa.z - (double)Pb.z) - uz;
              const double d2 = dx * dx + dy * dy + dz * dz;
              if (!(d2 < e2max)) continue;
              int lo = 0, hi = a.bins;
              while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (a.e2[mid] <= d2)
                  lo = mid;
                else
                  hi = mid;
              }
              atomicAdd(&hist[lo], 1u);
              ++cnt;
              sq += (unsigned long long)floor(sqrt_cr(d2) * 16777216.0 + 0.5);
              s2q += (unsigned long long)floor(d2 * 65536.0 + 0.5);
            }
          }
          __syncthreads();
        }
      }
    }
  }
  __syncthreads();
  for (int k = lane; k < a.bins; k += WAVE)
    if (hist[k]) atomicAdd(&a.out[k], (unsigned long long)hist[k]);
  cnt = wave_sum_u64(cnt);
  sq = wave_sum_u64(sq);
  s2q = wave_sum_u64(s2q);
  if (lane == 0) {
    atomicAdd(&a.out[a.bins + 0], cnt);
    atomicAdd(&a.out[a.bins + 1], sq);
    atomicAdd(&a.out[a.bins + 2], s2q);
  }
}

void launch_cent(const float4* P, const uint32_t* X, const uint32_t* start, const uint32_t* count,
                 const int4* st, uint32_t st_n, const int32_t nc[3], const double box[3],
                 const double* e2, int bins, float r2f, unsigned long long* out, hipStream_t s) {
  CentArg a;
  a.P = P;
  a.X = X;
  a.start = start;
  a.count = count;
  a.st = st;
  a.st_n = st_n;
  for (int d = 0; d < 3; ++d) {
    a.nc[d] = nc[d];
    a.box[d] = box[d];
  }
  a.e2 = e2;
  a.bins = bins;
  a.r2f = r2f;
  a.out = out;
  const uint64_t ncells = (uint64_t)nc[0] * nc[1] * nc[2];
  hipLaunchKernelGGL(k_cent, dim3(grid_for(ncells, 1)), dim3(WAVE), 0, s, a);
}

// ---------------------------------------------------------------------------
// Van der Waals energy over the sorted contact list (statKernel_VdW_naive,
// inc/statKernel_MUT.h:531-628, driven by collect_vdw,
// src/statistics_collection.cu:53-154): one thread per row sums its
// contributing pairs in row order; rows are summed in blocks of 1024 in
// order, block sums in order (deterministic).
// ---------------------------------------------------------------------------
__global__ void k_vdw_rows(const float4* __restrict__ P, const float4* __restrict__ D, uint32_t n,
                           const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ J,
                           const int32_t* __restrict__ img, double bx, double by, double bz,
                           double dlim, double A, double B, double* __restrict__ rowsum,
                           unsigned long long* __restrict__ pairs) {
  unsigned long long local = 0;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    double s = 0.0;
    const float4 P1 = P[r], D1 = D[r];
    for (uint32_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const int32_t c = img[k];
      const double sx = (double)(int)(int8_t)(c & 0xFF) * bx;
      const double sy = (double)(int)(int8_t)((c >> 8) & 0xFF) * by;
      const double sz = (double)(int)(int8_t)((c >> 16) & 0xFF) * bz;
      double e;
      if (vdw_pair(P1, D1, P[J[k]], D[J[k]], sx, sy, sz, dlim, A, B, &e)) {
        s = s + e;
        ++local;
      }
    }
    rowsum[r] = s;
  }
  local = wave_sum_u64(local);
  if ((threadIdx.x & (WAVE - 1)) == 0 && local) atomicAdd(pairs, local);
}

__global__ void k_block_sums(const double* __restrict__ v, uint64_t n, double* __restrict__ bsum) {
  const uint64_t nb = (n + 1023) / 1024;
  for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb;
       b += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = (b + 1) * 1024 < n ? (b + 1) * 1024 : n;
    double s = 0.0;
    for (uint64_t k = b * 1024; k < e; ++k) s = s + v[k];
    bsum[b] = s;
  }
}

__global__ void k_block_total(const double* __restrict__ bsum, uint64_t nb, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double t = 0.0;
  for (uint64_t b = 0; b < nb; ++b) t = t + bsum[b];
  *out = t;
}

void launch_vdw(const float4* P, const float4* D, uint32_t n, const uint32_t* rowptr,
                const uint32_t* J, const int32_t* img, const double box[3], double dlim, double A,
                double B, double* rowsum, double* bsum, double* out, unsigned long long* pairs,
                hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_vdw_rows, dim3(grid_for(n, 256)), dim3(256), 0, s, P, D, n, rowptr, J, img,
                       box[0], box[1], box[2], dlim, A, B, rowsum, pairs);
  const uint64_t nb = ((uint64_t)n + 1023) / 1024;
  if (nb > 0)
    hipLaunchKernelGGL(k_block_sums, dim3(grid_for(nb, 256)), dim3(256), 0, s, rowsum, (uint64_t)n,
                       bsum);
  hipLaunchKernelGGL(k_block_total, dim3(1), dim3(64), 0, s, bsum, nb, out);
}

// ---------------------------------------------------------------------------
// Histograms.  Angle bins by comparison with cos(bin edge): bin = number of
// edges k in 1..nb-1 with cos(angle) <= cos(edge_k), found by binary search
// on the decreasing edge table; the counts are exact (LDS u32 per block,
// then int64 atomics).  Replaces statKernel_ANGL (inc/statKernel_ANGL.h:
// 13-76) and statKernel_MUT_ANGLE_weightedFloat (inc/statKernel_MUT.h:
// 26-131) with their uint32/float accumulators.
// ---------------------------------------------------------------------------
constexpr int HIST_THREADS = 256;
constexpr int HIST_MAX_BINS = 256;
constexpr int HIST_MAX_2D = 4096;

__device__ __forceinline__ int edge_bin(double v, const double* __restrict__ e, int nb) {
  // largest k in [1, nb-1] with e[k] >= v, 0 if none (e decreasing)
  int lo = 0, hi = nb;  // answer in [lo, hi)
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (e[mid] >= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

struct HistArg {
  const double* ep;
  const double* em;
  const double* et;
  const double* ephi;
  int nb, npb, ntb, axis;
  double L;
  double mu[3];
};

__global__ void __launch_bounds__(HIST_THREADS) k_hist_rods(const float4* __restrict__ P,
                                                            const float4* __restrict__ D,
                                                            uint32_t n, HistArg h,
                                                            unsigned long long* __restrict__ out) {
  __shared__ uint32_t hc[HIST_MAX_BINS], hp[HIST_MAX_BINS];
  __shared__ uint32_t h2[HIST_MAX_2D];
  const int n2 = h.npb * h.ntb;
  for (int k = threadIdx.x; k < HIST_MAX_BINS; k += blockDim.x) hc[k] = hp[k] = 0;
  for (int k = threadIdx.x; k < n2; k += blockDim.x) h2[k] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = P[i], d = D[i];
    const float pa = h.axis == 0 ? p.x : (h.axis == 1 ? p.y : p.z);
    const double t = (double)pa / h.L;
    int b = (int)(t * (double)h.nb);
    b = b < 0 ? 0 : (b >= h.nb ? h.nb - 1 : b);
    atomicAdd(&hc[b], 1u);
    const double cx = d.x, cy = d.y, cz = d.z;
    const double cn = sqrt_cr((cx * cx + cy * cy) + cz * cz);
    const double cosv = ((cx * h.mu[0] + cy * h.mu[1]) + cz * h.mu[2]) / cn;
    atomicAdd(&hp[edge_bin(cosv, h.ep, h.nb)], 1u);
    if (n2 > 0) {
      const double ct = cz / cn;
      const int tb = edge_bin(ct, h.et, h.ntb);
      const double rxy = sqrt_cr(cx * cx + cy * cy);
      int pb = 0;
      if (rxy > 0.0) {
        const double cp = cx / rxy;
        const int bb = edge_bin(cp, h.ephi, h.npb / 2);
        pb = cy >= 0.0 ? bb : (h.npb - 1 - bb);
      }
      atomicAdd(&h2[tb * h.npb + pb], 1u);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < h.nb; k += blockDim.x) {
    if (hc[k]) atomicAdd(&out[k], (unsigned long long)hc[k]);
    if (hp[k]) atomicAdd(&out[h.nb + k], (unsigned long long)hp[k]);
  }
  for (int k = threadIdx.x; k < n2; k += blockDim.x)
    if (h2[k]) atomicAdd(&out[3 * h.nb + k], (unsigned long long)h2[k]);
}

// Mutual angle of contact pairs: one thread per row I of the contact CSR.
__global__ void __launch_bounds__(HIST_THREADS) k_hist_mut(const float4* __restrict__ D, uint32_t n,
                                                           const uint32_t* __restrict__ rowptr,
                                                           const uint32_t* __restrict__ J,
                                                           HistArg h,
                                                           unsigned long long* __restrict__ out) {
  __shared__ uint32_t hm[HIST_MAX_BINS];
  for (int k = threadIdx.x; k < HIST_MAX_BINS; k += blockDim.x) hm[k] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t b = rowptr[i], e = rowptr[i + 1];
    if (b == e) continue;
    const float4 a = D[i];
    const double ax = a.x, ay = a.y, az = a.z;
    const double na = sqrt_cr((ax * ax + ay * ay) + az * az);
    for (uint32_t k = b; k < e; ++k) {
      const float4 c = D[J[k]];
      const double bx = c.x, by = c.y, bz = c.z;
      const double nbv = sqrt_cr((bx * bx + by * by) + bz * bz);
      const double dv = fabs((ax * bx + ay * by) + az * bz) / (na * nbv);
      atomicAdd(&hm[edge_bin(dv, h.em, h.nb)], 1u);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < h.nb; k += blockDim.x)
    if (hm[k]) atomicAdd(&out[2 * h.nb + k], (unsigned long long)hm[k]);
}

// out layout: [centre nb][polar nb][mutual nb][phitheta npb*ntb]
void launch_hist(const float4* P, const float4* D, uint32_t n, const uint32_t* rowptr,
                 const uint32_t* J, const double* edges_p, const double* edges_m,
                 const double* edges_t, int nb, int npb, int ntb, int axis, double L,
                 const double* mu, unsigned long long* out, hipStream_t s) {
  HistArg h;
  h.ep = edges_p;
  h.em = edges_m;
  h.et = edges_t;
  h.ephi = edges_t ? edges_t + (ntb + 1) : nullptr;
  h.nb = nb;
  h.npb = npb;
  h.ntb = ntb;
  h.axis = axis;
  h.L = L;
  for (int a = 0; a < 3; ++a) h.mu[a] = mu[a];
  const int blocks = grid_for(n, HIST_THREADS) < 1024 ? grid_for(n, HIST_THREADS) : 1024;
  hipLaunchKernelGGL(k_hist_rods, dim3(blocks), dim3(HIST_THREADS), 0, s, P, D, n, h, out);
  if (rowptr)
    hipLaunchKernelGGL(k_hist_mut, dim3(blocks), dim3(HIST_THREADS), 0, s, D, n, rowptr, J, h,
                       out);
}

// ---------------------------------------------------------------------------
// Device arithmetic self-test: the FP64 building blocks whose results must be
// bit-identical to the CPU (correctly rounded sqrt, IEEE division, the
// library's exp/log, uniform conversion).
// ---------------------------------------------------------------------------
__global__ void k_selftest(int which, const double* __restrict__ in, double* __restrict__ out,
                           uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double r = 0.0;
    switch (which) {
      case 0: r = sqrt_cr(in[i]); break;
      case 1: r = dexp(in[i]); break;
      case 2: r = dlog(in[i]); break;
      case 3: r = in[2 * i] / in[2 * i + 1]; break;
      case 4: {
        const uint64_t b = __builtin_bit_cast(uint64_t, in[i]);
        r = u01((uint32_t)(b >> 32), (uint32_t)b);
        break;
      }
      case 6: r = in[i]; break;
      default: r = sqrt(in[i]); break;
    }
    out[i] = r;
  }
}

// The FP32 screen and band of one pair per thread, as the pair passes run
// them on the device: rods as 8 floats (centre, direction, length,
// radius), the FP64 image shift of rod b, the cut distance as the passes
// form it (contact: a_1 + a_2 + gap; generation: soft (a_1 + a_2) + sep).
__global__ void k_selftest_pairs(const float* __restrict__ ra, const float* __restrict__ rb,
                                 const double* __restrict__ sh, uint64_t n, int contact, double sep_or_gap,
                                 double soft_frac, float ew, int32_t* __restrict__ out_screen,
                                 int32_t* __restrict__ out_dec, int32_t* __restrict__ out_within) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const float* a = ra + 8 * i;
    const float* b = rb + 8 * i;
    const float4 P1 = make_float4(a[0], a[1], a[2], 0.5f * a[6]);
    const float4 D1 = make_float4(a[3], a[4], a[5], a[7]);
    const float4 P2 = make_float4(b[0], b[1], b[2], 0.5f * b[6]);
    const float4 D2 = make_float4(b[3], b[4], b[5], b[7]);
    const float R32 = contact ? ((D1.w + D2.w) + (float)sep_or_gap)
                              : ((float)soft_frac * (D1.w + D2.w) + (float)sep_or_gap);
    const float sx = (float)sh[3 * i], sy = (float)sh[3 * i + 1], sz = (float)sh[3 * i + 2];
    out_screen[i] = pair_f32_screen(P1, D1, P2, D2, sx, sy, sz, R32, ew);
    const F32Pair r = pair_f32(P1, D1, P2, D2, sx, sy, sz, R32, ew, !contact);
    out_dec[i] = r.decided;
    out_within[i] = r.within;
  }
}

void launch_selftest_pairs(const float* ra, const float* rb, const double* sh, uint64_t n, int contact,
                           double sep_or_gap, double soft_frac, float ew, int32_t* out_screen, int32_t* out_dec,
                           int32_t* out_within, hipStream_t s) {
  hipLaunchKernelGGL(k_selftest_pairs, dim3(grid_for(n, 256)), dim3(256), 0, s, ra, rb, sh, n, contact, sep_or_gap,
                     soft_frac, ew, out_screen, out_dec, out_within);
}

void launch_selftest(int which, const double* in, double* out, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_selftest, dim3(grid_for(n, 256)), dim3(256), 0, s, which, in, out, n);
}


// ---------------------------------------------------------------------------
// expand: periodic copies of the rods near the two faces normal to axis k.
// A rod whose centre lies within B of the far face (x >= L - B) gets a copy
// below the box, one within B of the near face (x < B) a copy above it;
// every coordinate along k is then shifted by B (new box [0, L + 2B)).  The
// shifts are FP64 sums rounded once to FP32.
// ---------------------------------------------------------------------------
__global__ void k_expand_flags(const float4* __restrict__ P, uint32_t n, int k, double L, double B,
                               uint32_t* __restrict__ below, uint32_t* __restrict__ above) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = P[i];
    const double x = (double)(k == 0 ? p.x : (k == 1 ? p.y : p.z));
    below[i] = x >= L - B ? 1u : 0u;
    above[i] = x < B ? 1u : 0u;
  }
}

__device__ __forceinline__ float4 expand_shift(float4 p, int k, double s) {
  if (k == 0)
    p.x = (float)((double)p.x + s);
  else if (k == 1)
    p.y = (float)((double)p.y + s);
  else
    p.z = (float)((double)p.z + s);
  return p;
}

__global__ void k_expand_scatter(const float4* __restrict__ P, const float4* __restrict__ D, uint32_t n, int k,
                                 double L, double B, const uint32_t* __restrict__ below,
                                 const uint32_t* __restrict__ rank_below, const uint32_t* __restrict__ above,
                                 const uint32_t* __restrict__ rank_above, uint32_t n_below,
                                 float4* __restrict__ oP, float4* __restrict__ oD) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = P[i], d = D[i];
    oP[i] = expand_shift(p, k, B);
    oD[i] = d;
    if (below[i]) {
      const uint32_t j = n + rank_below[i];
      oP[j] = expand_shift(p, k, B - L);
      oD[j] = d;
    }
    if (above[i]) {
      const uint32_t j = n + n_below + rank_above[i];
      oP[j] = expand_shift(p, k, B + L);
      oD[j] = d;
    }
  }
}

void launch_expand_flags(const float4* P, uint32_t n, int k, double L, double B, uint32_t* below, uint32_t* above,
                         hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_expand_flags, dim3(grid_for(n, 256)), dim3(256), 0, s, P, n, k, L, B, below, above);
}

void launch_expand_scatter(const float4* P, const float4* D, uint32_t n, int k, double L, double B,
                           const uint32_t* below, const uint32_t* rank_below, const uint32_t* above,
                           const uint32_t* rank_above, uint32_t n_below, float4* oP, float4* oD, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_expand_scatter, dim3(grid_for(n, 256)), dim3(256), 0, s, P, D, n, k, L, B, below,
                     rank_below, above, rank_above, n_below, oP, oD);
}

}  // namespace nf
