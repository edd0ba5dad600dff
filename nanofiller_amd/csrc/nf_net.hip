// nf_net.hip - network extraction on the contact graph of one packing:
// connected components (lock-free union-find, label = smallest rod id),
// electrode couplings and the FP64 Kirchhoff (graph Laplacian) matrix in CSR
// form on N + 2 nodes (rods 0..N-1, source electrode N, sink electrode N+1).
//
// The reference builds its contact network with Thrust sorts and partitions
// of 6-tuples (src/conductivity.cu:530-836, src/adjacency_graph.cu:57-219)
// and leaves the conductance matrix to an external Python solver
// (tools/kirchhoff/kirchhoff.py:211-242).  Here every row is produced in a
// canonical order - columns ascending, entries of one column (different
// periodic images) summed in image order, the diagonal summed over the
// merged entries in column order in 1024-entry blocks - so the matrix is
// bit-identical to a sequential evaluation of the same rules.
//
// Work is one thread per rod row: rows hold a handful of contacts, so the
// kernels are bound by HBM traffic over the contact arrays (about 20 bytes
// per contact and 40 bytes per rod per pass), not by arithmetic.
#include "nf_internal.h"

namespace nf {

namespace {

constexpr int NT = 256;

int grid_n(uint64_t n) {
  uint64_t b = (n + NT - 1) / NT;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

__device__ __forceinline__ int img_comp(int32_t code, int axis) {
  return (int)(int8_t)((code >> (8 * axis)) & 0xFF);
}

__device__ __forceinline__ int32_t img_neg(int32_t code) {
  const int x = -img_comp(code, 0), y = -img_comp(code, 1), z = -img_comp(code, 2);
  return (int32_t)(((uint32_t)(x & 0xFF)) | ((uint32_t)(y & 0xFF) << 8) |
                   ((uint32_t)(z & 0xFF) << 16));
}

__device__ __forceinline__ bool img_lt(int32_t a, int32_t b) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int x = img_comp(a, d), y = img_comp(b, d);
    if (x != y) return x < y;
  }
  return false;
}

// Tunnelling conductance g0 exp(-max(gap, 0) / lambda).
__device__ __forceinline__ double tunnel_g(double gap, double g0, double lam) {
  const double gg = gap > 0.0 ? gap : 0.0;
  return g0 * dexp(-(gg / lam));
}

// Sequential sum in 1024-entry blocks: block sums accumulate into tot.
struct BlockSum {
  double tot = 0.0, s = 0.0;
  uint32_t c = 0;
  __device__ void add(double v) {
    if (c == 1024u) {
      tot = tot + s;
      s = 0.0;
      c = 0;
    }
    s = s + v;
    ++c;
  }
  __device__ double result() const { return c ? tot + s : tot; }
};

// ---------------------------------------------------------------------------
// union-find.  Invariant: par[x] <= x, so every root is the smallest id of
// its tree and the final label of a component is its smallest rod id.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ld_par(const uint32_t* par, uint32_t x) {
  return __hip_atomic_load(&par[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ uint32_t uf_find(uint32_t* par, uint32_t x) {
  uint32_t p = ld_par(par, x);
  while (p != x) {
    const uint32_t gp = ld_par(par, p);
    // path halving: any ancestor of x is a valid parent, and the ids on the
    // path decrease, so atomicMin keeps the invariant
    if (gp != p) atomicMin(&par[x], gp);
    x = p;
    p = gp;
  }
  return x;
}

// Hooks the larger root under the smaller; a failed CAS means another
// thread re-hooked that root, so every retry follows progress elsewhere.
__device__ void uf_union(uint32_t* par, uint32_t a, uint32_t b) {
  for (;;) {
    const uint32_t ra = uf_find(par, a), rb = uf_find(par, b);
    if (ra == rb) return;
    const uint32_t hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
    if (atomicCAS(&par[hi], hi, lo) == hi) return;
  }
}

__global__ void k_net_init(NetArgs a) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
    a.par[i] = i;
    a.lowcnt[i] = 0u;
    a.lowfill[i] = 0u;
    a.csize[i] = 0u;
    a.cflag[i] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x < NET_SUMMARY) a.summary[threadIdx.x] = 0ull;
}

// Rod rows: union over the network edges (contacts that do not cross the
// electrode-axis boundary), lower-triangle entry counts per column.
__global__ void k_net_edges(NetArgs a) {
  unsigned long long local = 0;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.n; r += gridDim.x * blockDim.x) {
    const uint32_t e = a.rowptr[r + 1];
    for (uint32_t k = a.rowptr[r]; k < e; ++k) {
      if (img_comp(a.img[k], a.axis) != 0) continue;
      const uint32_t j = a.J[k];
      ++local;
      atomicAdd(&a.lowcnt[j], 1u);
      uf_union(a.par, r, j);
    }
  }
  // one atomic per wave, after the loop (all lanes converged)
  local = wave_sum_u64(local);
  if ((threadIdx.x & (WAVE - 1)) == 0 && local) atomicAdd(&a.summary[SUM_EDGES], local);
}

__global__ void k_net_flatten(NetArgs a) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x)
    atomicMin(&a.par[i], uf_find(a.par, i));
}

// Electrode couplings: the capsule's extent along the axis against the
// planes at 0 and L; coupled when the surface gap is within the cut.
__global__ void k_net_elec(NetArgs a) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
    const float4 p = a.P[i], d = a.D[i];
    const float pa = a.axis == 0 ? p.x : (a.axis == 1 ? p.y : p.z);
    const float ca = a.axis == 0 ? d.x : (a.axis == 1 ? d.y : d.z);
    const double ext = fabs((double)p.w * (double)ca);  // (l * 0.5) * c; l * 0.5 is exact
    const double pos = (double)pa, rad = (double)d.w;
    const double gs = (pos - ext) - rad;
    const double gk = (a.L - (pos + ext)) - rad;
    const bool cs = gs <= a.gcut, ck = gk <= a.gcut;
    a.fsrc[i] = cs ? 1u : 0u;
    a.fsnk[i] = ck ? 1u : 0u;
    a.gsrc[i] = cs ? tunnel_g(gs, a.g0, a.lam) : 0.0;
    a.gsnk[i] = ck ? tunnel_g(gk, a.g0, a.lam) : 0.0;
  }
}

__global__ void k_net_comp(NetArgs a) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
    const uint32_t r = a.par[i];
    atomicAdd(&a.csize[r], 1u);
    const uint32_t f = (a.fsrc[i] ? 1u : 0u) | (a.fsnk[i] ? 2u : 0u);
    if (f) atomicOr(&a.cflag[r], f);
  }
}

__global__ void k_net_info(NetArgs a) {
  unsigned long long comps = 0, big = 0, perc = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
    if (a.par[i] != i) continue;
    ++comps;
    big = a.csize[i] > big ? a.csize[i] : big;
    perc += a.cflag[i] == 3u ? 1u : 0u;
  }
  comps = wave_sum_u64(comps);
  big = wave_max_u64(big);
  perc = wave_sum_u64(perc);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (comps) atomicAdd(&a.summary[SUM_COMPONENTS], comps);
    if (big) atomicMax(&a.summary[SUM_LARGEST], big);
    if (perc) atomicAdd(&a.summary[SUM_PERCOLATING], perc);
  }
}

// Lower-triangle entries of row j gathered from the contact rows r < j,
// with the image seen from j (negated).
__global__ void k_net_lower_fill(NetArgs a) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.n; r += gridDim.x * blockDim.x) {
    const uint32_t e = a.rowptr[r + 1];
    for (uint32_t k = a.rowptr[r]; k < e; ++k) {
      const int32_t im = a.img[k];
      if (img_comp(im, a.axis) != 0) continue;
      const uint32_t j = a.J[k];
      const uint32_t pos = a.lowstart[j] + atomicAdd(&a.lowfill[j], 1u);
      a.lcol[pos] = r;
      a.limg[pos] = img_neg(im);
      a.lg[pos] = tunnel_g(a.gap64[k], a.g0, a.lam);
    }
  }
}

// Sorts each row's lower entries by (column, image); (column, image) pairs
// are unique, so the result does not depend on the fill order.
__global__ void k_net_lower_sort(NetArgs a) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.n; r += gridDim.x * blockDim.x) {
    const uint32_t b = a.lowstart[r], e = a.lowstart[r + 1];
    for (uint32_t i = b + 1; i < e; ++i) {
      const uint32_t kc = a.lcol[i];
      const int32_t ki = a.limg[i];
      const double kg = a.lg[i];
      uint32_t k = i;
      while (k > b && (a.lcol[k - 1] > kc || (a.lcol[k - 1] == kc && img_lt(ki, a.limg[k - 1])))) {
        a.lcol[k] = a.lcol[k - 1];
        a.limg[k] = a.limg[k - 1];
        a.lg[k] = a.lg[k - 1];
        --k;
      }
      a.lcol[k] = kc;
      a.limg[k] = ki;
      a.lg[k] = kg;
    }
  }
}

// Receives the merged entries of one rod row in column order and keeps the
// diagonal's blocked sum.
struct RowSink {
  uint32_t* col;
  double* val;
  uint32_t pos;
  uint32_t count;
  BlockSum diag;
  template <bool WRITE>
  __device__ void put(uint32_t c, double g) {
    diag.add(g);
    if (WRITE) {
      col[pos] = c;
      val[pos] = -g;
      ++pos;
    }
    ++count;
  }
};

// Walks row r: merged lower entries, the diagonal slot, merged upper
// entries (the contact row, sorted by column then image), source, sink.
template <bool WRITE>
__device__ void row_walk(const NetArgs& a, uint32_t r, RowSink& o) {
  {
    uint32_t prev = 0xFFFFFFFFu;
    double acc = 0.0;
    const uint32_t e = a.lowstart[r + 1];
    for (uint32_t k = a.lowstart[r]; k < e; ++k) {
      const uint32_t c = a.lcol[k];
      if (c == prev) {
        acc = acc + a.lg[k];
      } else {
        if (prev != 0xFFFFFFFFu) o.template put<WRITE>(prev, acc);
        prev = c;
        acc = a.lg[k];
      }
    }
    if (prev != 0xFFFFFFFFu) o.template put<WRITE>(prev, acc);
  }
  const uint32_t dpos = o.pos;
  if (WRITE) ++o.pos;
  {
    uint32_t prev = 0xFFFFFFFFu;
    double acc = 0.0;
    const uint32_t e = a.rowptr[r + 1];
    for (uint32_t k = a.rowptr[r]; k < e; ++k) {
      if (img_comp(a.img[k], a.axis) != 0) continue;
      const uint32_t c = a.J[k];
      const double g = tunnel_g(a.gap64[k], a.g0, a.lam);
      if (c == prev) {
        acc = acc + g;
      } else {
        if (prev != 0xFFFFFFFFu) o.template put<WRITE>(prev, acc);
        prev = c;
        acc = g;
      }
    }
    if (prev != 0xFFFFFFFFu) o.template put<WRITE>(prev, acc);
  }
  if (a.fsrc[r]) o.template put<WRITE>(a.n, a.gsrc[r]);
  if (a.fsnk[r]) o.template put<WRITE>(a.n + 1, a.gsnk[r]);
  if (WRITE) {
    o.col[dpos] = r;
    o.val[dpos] = o.diag.result();
  }
}

__global__ void k_net_rowcount(NetArgs a) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.n; r += gridDim.x * blockDim.x) {
    RowSink o{nullptr, nullptr, 0u, 0u, BlockSum()};
    row_walk<false>(a, r, o);
    a.rowcnt[r] = o.count + 1u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.rowcnt[a.n] = a.psrc[a.n] + 1u;
    a.rowcnt[a.n + 1] = a.psnk[a.n] + 1u;
  }
}

__global__ void k_net_rowfill(NetArgs a) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.n; r += gridDim.x * blockDim.x) {
    RowSink o{a.ncol, a.nval, a.nrow[r], 0u, BlockSum()};
    row_walk<true>(a, r, o);
  }
}

__global__ void k_net_summary(NetArgs a) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  a.summary[SUM_NNZ] = a.nrow[a.n + 2];
  a.summary[SUM_SOURCE] = a.psrc[a.n];
  a.summary[SUM_SINK] = a.psnk[a.n];
}

// Electrode row: the coupled rods in id order (positions from the scan of
// the coupling flags); their conductances are compacted into cg for the
// diagonal's blocked sum.
__global__ void k_net_elec_row(NetArgs a, int sink) {
  const uint32_t* flag = sink ? a.fsnk : a.fsrc;
  const uint32_t* fpos = sink ? a.psnk : a.psrc;
  const double* g = sink ? a.gsnk : a.gsrc;
  const uint32_t base = a.nrow[a.n + (sink ? 1 : 0)];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
    if (!flag[i]) continue;
    const uint32_t q = fpos[i];
    a.ncol[base + q] = i;
    a.nval[base + q] = -g[i];
    a.cg[q] = g[i];
  }
}

__global__ void k_net_elec_blocks(NetArgs a, int sink) {
  const uint32_t cnt = (sink ? a.psnk : a.psrc)[a.n];
  const uint32_t nb = (cnt + 1023u) / 1024u;
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
    const uint32_t e = (b + 1u) * 1024u < cnt ? (b + 1u) * 1024u : cnt;
    double s = 0.0;
    for (uint32_t k = b * 1024u; k < e; ++k) s = s + a.cg[k];
    a.bsum[b] = s;
  }
}

__global__ void k_net_elec_diag(NetArgs a, int sink) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const uint32_t cnt = (sink ? a.psnk : a.psrc)[a.n];
  const uint32_t nb = (cnt + 1023u) / 1024u;
  double tot = 0.0;
  for (uint32_t b = 0; b < nb; ++b) tot = tot + a.bsum[b];
  const uint32_t pos = a.nrow[a.n + (sink ? 1 : 0)] + cnt;
  a.ncol[pos] = a.n + (sink ? 1u : 0u);
  a.nval[pos] = tot;
}

}  // namespace

void launch_network_build(const NetArgs& a, hipStream_t s) {
  const int g = grid_n(a.n);
  hipLaunchKernelGGL(k_net_init, dim3(g), dim3(NT), 0, s, a);
  hipLaunchKernelGGL(k_net_edges, dim3(g), dim3(NT), 0, s, a);
  hipLaunchKernelGGL(k_net_flatten, dim3(g), dim3(NT), 0, s, a);
  hipLaunchKernelGGL(k_net_elec, dim3(g), dim3(NT), 0, s, a);
  hipLaunchKernelGGL(k_net_comp, dim3(g), dim3(NT), 0, s, a);
  hipLaunchKernelGGL(k_net_info, dim3(g), dim3(NT), 0, s, a);
  scan_u32(a.lowcnt, a.lowstart, a.n, a.tmp, s);
  hipLaunchKernelGGL(k_net_lower_fill, dim3(g), dim3(NT), 0, s, a);
  hipLaunchKernelGGL(k_net_lower_sort, dim3(g), dim3(NT), 0, s, a);
  scan_u32(a.fsrc, a.psrc, a.n, a.tmp, s);
  scan_u32(a.fsnk, a.psnk, a.n, a.tmp, s);
  hipLaunchKernelGGL(k_net_rowcount, dim3(g), dim3(NT), 0, s, a);
  scan_u32(a.rowcnt, a.nrow, (uint64_t)a.n + 2, a.tmp, s);
  hipLaunchKernelGGL(k_net_summary, dim3(1), dim3(64), 0, s, a);
}

void launch_network_fill(const NetArgs& a, uint32_t n_src, uint32_t n_snk, hipStream_t s) {
  const int g = grid_n(a.n);
  hipLaunchKernelGGL(k_net_rowfill, dim3(g), dim3(NT), 0, s, a);
  for (int sink = 0; sink < 2; ++sink) {
    const uint32_t cnt = sink ? n_snk : n_src;
    hipLaunchKernelGGL(k_net_elec_row, dim3(g), dim3(NT), 0, s, a, sink);
    hipLaunchKernelGGL(k_net_elec_blocks, dim3(grid_n((cnt + 1023u) / 1024u)), dim3(NT), 0, s, a,
                       sink);
    hipLaunchKernelGGL(k_net_elec_diag, dim3(1), dim3(64), 0, s, a, sink);
  }
}

}  // namespace nf
