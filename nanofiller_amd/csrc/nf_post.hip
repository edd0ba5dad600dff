// nf_post.hip - cell-list build for a given packing, canonical ordering of
// contact rows, and exact int64 histograms.
#include "nf_internal.h"

namespace nf {

__device__ __forceinline__ uint32_t cell_index(const float4 P, const int32_t nc[3],
                                               const double sc[3]) {
  const float p[3] = {P.x, P.y, P.z};
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int i = (int)((double)p[a] * sc[a]);
    c[a] = i < 0 ? 0 : (i >= nc[a] ? nc[a] - 1 : i);
  }
  return (uint32_t)(c[0] + nc[0] * (c[1] + nc[1] * c[2]));
}

struct GridArg {
  int32_t nc[3];
  double sc[3];
};

// Cell binning of a packing (count pass): the role of writeGlbMem's
// per-cell placement (inc/memoryIO.h:361-429) for an existing packing.
__global__ void k_cell_count(const float4* __restrict__ P, uint32_t n, GridArg g,
                             uint32_t* __restrict__ count, float* __restrict__ hl) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = P[i];
    const uint32_t c = cell_index(p, g.nc, g.sc);
    atomicAdd(&count[c], 1u);
    atomicMax((unsigned int*)&hl[c], __float_as_uint(p.w));
  }
}

__global__ void k_cell_scatter(const float4* __restrict__ P, const float4* __restrict__ D, uint32_t n,
                               GridArg g, const uint32_t* __restrict__ start,
                               uint32_t* __restrict__ cur, float4* __restrict__ oP,
                               float4* __restrict__ oD, uint32_t* __restrict__ oid) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 p = P[i];
    const uint32_t c = cell_index(p, g.nc, g.sc);
    const uint32_t dst = start[c] + atomicAdd(&cur[c], 1u);
    oP[dst] = p;
    oD[dst] = D[i];
    oid[dst] = i;
  }
}

static int grid_for(uint64_t n, int threads) {
  uint64_t b = (n + threads - 1) / threads;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

void launch_cell_count(const float4* P, uint32_t n, const int32_t ncell[3], const double scale[3],
                       uint32_t* count, float* hl, hipStream_t s) {
  GridArg g;
  for (int a = 0; a < 3; ++a) {
    g.nc[a] = ncell[a];
    g.sc[a] = scale[a];
  }
  hipLaunchKernelGGL(k_cell_count, dim3(grid_for(n, 256)), dim3(256), 0, s, P, n, g, count, hl);
}

void launch_cell_scatter(const float4* P, const float4* D, uint32_t n, const int32_t ncell[3],
                         const double scale[3], const uint32_t* start, uint32_t* cur, float4* oP,
                         float4* oD, uint32_t* oid, hipStream_t s) {
  GridArg g;
  for (int a = 0; a < 3; ++a) {
    g.nc[a] = ncell[a];
    g.sc[a] = scale[a];
  }
  hipLaunchKernelGGL(k_cell_scatter, dim3(grid_for(n, 256)), dim3(256), 0, s, P, D, n, g, start,
                     cur, oP, oD, oid);
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

void launch_selftest(int which, const double* in, double* out, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_selftest, dim3(grid_for(n, 256)), dim3(256), 0, s, which, in, out, n);
}

}  // namespace nf
