// nf_solve.hip - DC effective conductance of the Kirchhoff network between
// the two electrodes: source at potential 1, sink at 0.
//
// Replaces the solve of tools/kirchhoff/kirchhoff.py:256-272 (scipy
// spsolve or MKL DSS on diag(S) - A - A^T with the source couplings as
// right-hand side; current I = sum (1 - V_i) B_i) for the real, DC
// conductances g0 exp(-gap / lambda) of the network matrix.
//
// Only components that touch both electrodes carry current: a component
// touching one electrode sits at its potential, one touching none is
// floating.  The rods of the current-carrying components are compacted
// into a reduced system A x = b (A: the Laplacian rows restricted to those
// rods, electrode couplings kept on the diagonal; b: the source couplings),
// which is symmetric positive definite, and solved by Jacobi-preconditioned
// conjugate gradients in FP64.
//
// Every reduction is a fixed-order tree over fixed per-block ranges, so a
// solve is bit-reproducible on one device and grid.  The kernels stream the
// reduced CSR and five vectors per iteration and are bound by HBM bandwidth
// (or, on small systems, by launch latency).
#include "nf_internal.h"

namespace nf {

namespace {

constexpr int NT = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  // wave sums in a fixed butterfly order, then the waves' sums in order
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  const int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < NT / WAVE; ++k) t += sh[k];
  __syncthreads();
  return t;  // valid in thread 0
}

// Marks the components coupled to the source (bit 0) and the sink (bit 1):
// the electrode rows N and N+1 list their coupled rods.
__global__ void __launch_bounds__(NT) k_cs_flag(const uint32_t* __restrict__ nrow,
                                                const uint32_t* __restrict__ ncol,
                                                const uint32_t* __restrict__ labels, uint32_t n,
                                                uint32_t* __restrict__ flag) {
  const uint32_t b = nrow[n], e = nrow[n + 2], mid = nrow[n + 1];
  for (uint32_t k = b + blockIdx.x * NT + threadIdx.x; k < e; k += gridDim.x * NT) {
    const uint32_t j = ncol[k];
    if (j < n) atomicOr(&flag[labels[j]], k < mid ? 1u : 2u);
  }
}

// act[i] = 1 for rods in components touching both electrodes; rl[i] = the
// number of rod columns in their rows (0 elsewhere).
__global__ void __launch_bounds__(NT) k_cs_active(const uint32_t* __restrict__ nrow,
                                                  const uint32_t* __restrict__ ncol,
                                                  const uint32_t* __restrict__ labels,
                                                  const uint32_t* __restrict__ flag, uint32_t n,
                                                  uint32_t* __restrict__ act,
                                                  uint32_t* __restrict__ rl) {
  for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
    const bool a = flag[labels[i]] == 3u;
    uint32_t c = 0;
    if (a)
      for (uint32_t k = nrow[i]; k < nrow[i + 1]; ++k) c += ncol[k] < n ? 1u : 0u;
    act[i] = a ? 1u : 0u;
    rl[i] = c;
  }
}

// Reduced system: CSR (columns renumbered, ascending order kept), inverse
// diagonal, source and sink couplings.
__global__ void __launch_bounds__(NT) k_cs_fill(const uint32_t* __restrict__ nrow,
                                                const uint32_t* __restrict__ ncol,
                                                const double* __restrict__ nval, uint32_t n,
                                                const uint32_t* __restrict__ act,
                                                const uint32_t* __restrict__ idx,
                                                const uint32_t* __restrict__ off,
                                                uint32_t* __restrict__ crow,
                                                uint32_t* __restrict__ ccol,
                                                double* __restrict__ cval,
                                                double* __restrict__ dinv, double* __restrict__ bsrc,
                                                double* __restrict__ bsnk) {
  for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
    if (!act[i]) continue;
    const uint32_t r = idx[i];
    uint32_t pos = off[i];
    crow[r] = pos;
    double gs = 0.0, gk = 0.0, dg = 1.0;
    for (uint32_t k = nrow[i]; k < nrow[i + 1]; ++k) {
      const uint32_t j = ncol[k];
      const double v = nval[k];
      if (j < n) {
        ccol[pos] = idx[j];
        cval[pos] = v;
        ++pos;
        if (j == i) dg = v;
      } else if (j == n) {
        gs = -v;
      } else {
        gk = -v;
      }
    }
    dinv[r] = 1.0 / dg;
    bsrc[r] = gs;
    bsnk[r] = gk;
  }
}

// Unknown numbering in a spatial order: ord lists the rods by cell, so
// rods that share contacts get nearby unknowns and the matrix-vector
// product's gathers stay within cached lines.  Gathers the active flags
// and row lengths into that order; the scans of them are scattered back to
// per-rod unknown indices and row offsets.
__global__ void __launch_bounds__(NT) k_cs_to_order(uint32_t n, const uint32_t* __restrict__ ord,
                                                    const uint32_t* __restrict__ act,
                                                    const uint32_t* __restrict__ rl, uint32_t* __restrict__ acto,
                                                    uint32_t* __restrict__ rlo) {
  for (uint32_t k = blockIdx.x * NT + threadIdx.x; k < n; k += gridDim.x * NT) {
    const uint32_t i = ord[k];
    acto[k] = act[i];
    rlo[k] = rl[i];
  }
}

__global__ void __launch_bounds__(NT) k_cs_from_order(uint32_t n, const uint32_t* __restrict__ ord,
                                                      const uint32_t* __restrict__ idxo,
                                                      const uint32_t* __restrict__ offo,
                                                      uint32_t* __restrict__ idx, uint32_t* __restrict__ off) {
  for (uint32_t k = blockIdx.x * NT + threadIdx.x; k <= n; k += gridDim.x * NT) {
    const uint32_t i = k < n ? ord[k] : n;  // the totals stay at n
    idx[i] = idxo[k];
    off[i] = offo[k];
  }
}

// x = 0, r = b, z = D^-1 r, p = z; partials of r.z and b.b.
__global__ void __launch_bounds__(NT) k_cg_init(uint32_t m, const double* __restrict__ b,
                                                const double* __restrict__ dinv,
                                                double* __restrict__ x, double* __restrict__ r,
                                                double* __restrict__ z, double* __restrict__ p,
                                                double* __restrict__ part) {
  __shared__ double sh[NT / WAVE];
  double rz = 0.0, bb = 0.0;
  for (uint32_t k = blockIdx.x * NT + threadIdx.x; k < m; k += gridDim.x * NT) {
    const double bk = b[k], zk = dinv[k] * bk;
    x[k] = 0.0;
    r[k] = bk;
    z[k] = zk;
    p[k] = zk;
    rz += bk * zk;
    bb += bk * bk;
  }
  const double s1 = block_sum(rz, sh);
  const double s2 = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

// q = A p; partial of p.q.  Four lanes per row: lane j of a row sums its
// entries j, j + 4, ... (consecutive lanes read consecutive entries, so the
// column and value loads coalesce across the rows of a wave), then the four
// lane sums combine in a fixed butterfly, so the result is reproducible.
constexpr int SPMV_LANES = 4;

__global__ void __launch_bounds__(NT) k_cg_spmv(uint32_t m, const uint32_t* __restrict__ crow,
                                                const uint32_t* __restrict__ ccol,
                                                const double* __restrict__ cval,
                                                const double* __restrict__ p, double* __restrict__ q,
                                                double* __restrict__ part) {
  __shared__ double sh[NT / WAVE];
  const uint32_t sub = threadIdx.x & (SPMV_LANES - 1);
  const uint32_t rows_per_block = NT / SPMV_LANES;
  double pq = 0.0;
  // block-uniform loop over row groups, so the shuffles run with every lane
  for (uint32_t k0 = blockIdx.x * rows_per_block; k0 < m; k0 += gridDim.x * rows_per_block) {
    const uint32_t k = k0 + threadIdx.x / SPMV_LANES;
    double s = 0.0;
    if (k < m) {
      const uint32_t e1 = crow[k + 1];
      for (uint32_t e = crow[k] + sub; e < e1; e += SPMV_LANES) s += cval[e] * p[ccol[e]];
    }
    s += __shfl_xor(s, 1, WAVE);
    s += __shfl_xor(s, 2, WAVE);
    if (k < m && sub == 0u) {
      q[k] = s;
      pq += p[k] * s;
    }
  }
  const double t = block_sum(pq, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = t;
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// One block: sums the nb partial pairs in block order.  which = 0: init
// (rz, bb); 1: alpha = rz / pq; 2: beta = rz' / rz, rz = rz', rr;
// 3: source and sink currents.
__global__ void __launch_bounds__(NT) k_cg_scalar(const double* __restrict__ part, int nb, int which,
                                                  CgScal* __restrict__ sc) {
  __shared__ double sh[NT / WAVE];
  double a = 0.0, c = 0.0;
  for (int k = threadIdx.x; k < nb; k += NT) {
    a += part[2 * k];
    c += part[2 * k + 1];
  }
  const double sa = block_sum(a, sh);
  const double sb = block_sum(c, sh);
  if (threadIdx.x != 0) return;
  if (which == 0) {
    sc->rz = sa;
    sc->bb = sb;
    sc->rr = sb;
  } else if (which == 1) {
    sc->pq = sa;
    sc->alpha = sa != 0.0 ? sc->rz / sa : 0.0;
  } else if (which == 3) {
    sc->isrc = sa;
    sc->isnk = sb;
  } else {
    sc->beta = sc->rz != 0.0 ? sa / sc->rz : 0.0;
    sc->rz = sa;
    sc->rr = sb;
  }
}

// x += alpha p, r -= alpha q, z = D^-1 r; partials of r.z and r.r.
__global__ void __launch_bounds__(NT) k_cg_update(uint32_t m, const CgScal* __restrict__ sc,
                                                  const double* __restrict__ dinv,
                                                  const double* __restrict__ p,
                                                  const double* __restrict__ q,
                                                  double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ z,
                                                  double* __restrict__ part) {
  __shared__ double sh[NT / WAVE];
  const double alpha = sc->alpha;
  double rz = 0.0, rr = 0.0;
  for (uint32_t k = blockIdx.x * NT + threadIdx.x; k < m; k += gridDim.x * NT) {
    x[k] += alpha * p[k];
    const double rk = r[k] - alpha * q[k];
    const double zk = dinv[k] * rk;
    r[k] = rk;
    z[k] = zk;
    rz += rk * zk;
    rr += rk * rk;
  }
  const double s1 = block_sum(rz, sh);
  const double s2 = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

// p = z + beta p.
__global__ void __launch_bounds__(NT) k_cg_dir(uint32_t m, const CgScal* __restrict__ sc,
                                               const double* __restrict__ z,
                                               double* __restrict__ p) {
  const double beta = sc->beta;
  for (uint32_t k = blockIdx.x * NT + threadIdx.x; k < m; k += gridDim.x * NT) p[k] = z[k] + beta * p[k];
}

// Partials of the source current sum bs (1 - x) and the sink current sum bk x.
__global__ void __launch_bounds__(NT) k_cg_current(uint32_t m, const double* __restrict__ bs,
                                                   const double* __restrict__ bk,
                                                   const double* __restrict__ x,
                                                   double* __restrict__ part) {
  __shared__ double sh[NT / WAVE];
  double is = 0.0, ik = 0.0;
  for (uint32_t k = blockIdx.x * NT + threadIdx.x; k < m; k += gridDim.x * NT) {
    is += bs[k] * (1.0 - x[k]);
    ik += bk[k] * x[k];
  }
  const double s1 = block_sum(is, sh);
  const double s2 = block_sum(ik, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1;
    part[2 * blockIdx.x + 1] = s2;
  }
}

// Potentials of every rod: solved, electrode potential, or NaN (floating).
__global__ void __launch_bounds__(NT) k_cs_potential(uint32_t n, const uint32_t* __restrict__ labels,
                                                     const uint32_t* __restrict__ flag,
                                                     const uint32_t* __restrict__ act,
                                                     const uint32_t* __restrict__ idx,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ v) {
  for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < n; i += gridDim.x * NT) {
    const uint32_t f = flag[labels[i]];
    double vi = __longlong_as_double(0x7FF8000000000000ll);
    if (act[i])
      vi = x[idx[i]];
    else if (f == 1u)
      vi = 1.0;
    else if (f == 2u)
      vi = 0.0;
    v[i] = vi;
  }
}

int grid_for(uint64_t n, int cap) {
  uint64_t b = (n + NT - 1) / NT;
  if (b > (uint64_t)cap) b = cap;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

int cg_grid(uint64_t m) { return grid_for(m, CG_MAX_BLOCKS); }

void launch_cs_flag(const uint32_t* nrow, const uint32_t* ncol, const uint32_t* labels, uint32_t n,
                    uint32_t* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_cs_flag, dim3(64), dim3(NT), 0, s, nrow, ncol, labels, n, flag);
}

void launch_cs_active(const uint32_t* nrow, const uint32_t* ncol, const uint32_t* labels,
                      const uint32_t* flag, uint32_t n, uint32_t* act, uint32_t* rl, hipStream_t s) {
  hipLaunchKernelGGL(k_cs_active, dim3(grid_for(n, 8192)), dim3(NT), 0, s, nrow, ncol, labels, flag,
                     n, act, rl);
}

void launch_cs_fill(const uint32_t* nrow, const uint32_t* ncol, const double* nval, uint32_t n,
                    const uint32_t* act, const uint32_t* idx, const uint32_t* off, uint32_t* crow,
                    uint32_t* ccol, double* cval, double* dinv, double* bsrc, double* bsnk,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_cs_fill, dim3(grid_for(n, 8192)), dim3(NT), 0, s, nrow, ncol, nval, n, act,
                     idx, off, crow, ccol, cval, dinv, bsrc, bsnk);
}

void launch_cs_potential(uint32_t n, const uint32_t* labels, const uint32_t* flag,
                         const uint32_t* act, const uint32_t* idx, const double* x, double* v,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_cs_potential, dim3(grid_for(n, 8192)), dim3(NT), 0, s, n, labels, flag, act,
                     idx, x, v);
}

void launch_cs_order(uint32_t n, const uint32_t* ord, const uint32_t* act, const uint32_t* rl, uint32_t* acto,
                     uint32_t* rlo, const uint32_t* idxo, const uint32_t* offo, uint32_t* idx, uint32_t* off,
                     int stage, hipStream_t s) {
  if (stage == 0)
    hipLaunchKernelGGL(k_cs_to_order, dim3(grid_for(n, 8192)), dim3(NT), 0, s, n, ord, act, rl, acto, rlo);
  else
    hipLaunchKernelGGL(k_cs_from_order, dim3(grid_for((uint64_t)n + 1, 8192)), dim3(NT), 0, s, n, ord, idxo, offo,
                       idx, off);
}

void launch_cg_init(const CgArgs& a, hipStream_t s) {
  const int g = cg_grid(a.m);
  hipLaunchKernelGGL(k_cg_init, dim3(g), dim3(NT), 0, s, a.m, a.bsrc, a.dinv, a.x, a.r, a.z, a.p,
                     a.part);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(NT), 0, s, a.part, g, 0, a.scal);
}

void launch_cg_iteration(const CgArgs& a, hipStream_t s) {
  const int g = cg_grid(a.m);
  CgScal* sc = a.scal;
  hipLaunchKernelGGL(k_cg_spmv, dim3(g), dim3(NT), 0, s, a.m, a.crow, a.ccol, a.cval, a.p, a.q,
                     a.part);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(NT), 0, s, a.part, g, 1, sc);
  hipLaunchKernelGGL(k_cg_update, dim3(g), dim3(NT), 0, s, a.m, sc, a.dinv, a.p, a.q, a.x, a.r, a.z,
                     a.part);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(NT), 0, s, a.part, g, 2, sc);
  hipLaunchKernelGGL(k_cg_dir, dim3(g), dim3(NT), 0, s, a.m, sc, a.z, a.p);
}

void launch_cg_current(const CgArgs& a, hipStream_t s) {
  const int g = cg_grid(a.m);
  hipLaunchKernelGGL(k_cg_current, dim3(g), dim3(NT), 0, s, a.m, a.bsrc, a.bsnk, a.x, a.part);
  hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(NT), 0, s, a.part, g, 3, a.scal);
}

}  // namespace nf
