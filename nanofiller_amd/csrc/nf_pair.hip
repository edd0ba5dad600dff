// nf_pair.hip - the cell-pair engine.
//
// One 64-lane wavefront (one workgroup) per home cell, taken from an
// XCD-aware work queue.  The home cell's probes sit in registers; each
// stencil neighbour cell is streamed through LDS in 64-rod tiles.  When a
// cell holds G < 64 probes, K = 64/next_pow2(G) lanes share a probe and
// stride the tile, so a sparsely filled cell still keeps the wave busy.
// Stencil offsets are sorted by the minimum cell-to-cell distance, so the
// loop stops at the first offset beyond the chunk's reach; cells whose
// longest rod cannot reach the probes are skipped.
//
// Modes replace the reference's kernels:
//   PM_GEN_OLD    new vs accepted rods   cuda_OldExtIsectLB (inc/intersection_kern.h:245-364)
//   PM_GEN_NEW    new vs new, priority   cuda_InternalIsect_noLck / cuda_NewExtIsectLB_parted
//                                        (inc/intersection_kern.h:12-241)
//   PM_CNT_COUNT  contacts per rod       core_cell_interact_kernel CONTACT_COUNT
//   PM_CNT_RECORD contact records        core_cell_interact_kernel CONTACT_RECORD
//                                        (src/core_cell_interaction_kernel.cu:70-315)
#include "nf_internal.h"

namespace nf {

constexpr int PAIR_THREADS = WAVE;

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, WAVE));
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

__device__ __forceinline__ int32_t pack_img(int x, int y, int z) {
  return (int32_t)(((uint32_t)(x & 0xFF)) | ((uint32_t)(y & 0xFF) << 8) |
                   ((uint32_t)(z & 0xFF) << 16));
}

struct CellGeom {
  double box[3];
  float boxf[3];
  int nc[3];
  uint32_t cell_base;
  const int4* st;
  uint32_t st_n;
  float ew, hl_max, R_max;
  double sf64, sep64;
  float sf32, sep32;
  uint32_t cap;
  InstState* state;
};

// Per-lane statistics, flushed to the instance (or the contact counters)
// by every lane together; lanes other than 0 add zero.
struct Acc {
  uint32_t tests, rech, kills;
  double maxdd;
  unsigned long long bytes;  // wave-uniform: probe and tile bytes streamed
};

// Values the hardware must treat as wave-uniform: read from lane 0 into an
// SGPR so that every branch and loop bound that depends on them is scalar.
// The pair engine keeps all control flow outside the innermost tile loop
// uniform in this way, and performs cross-lane operations (shuffles,
// ballots, barriers) only where every lane of the wave is active.
__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ int unii(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float unif(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

template <bool GEN>
__device__ __forceinline__ void flush(const PairArgs& a, const CellGeom& g, Acc& acc, int mode) {
  const uint32_t wt = wave_sum(acc.tests), wr = wave_sum(acc.rech), wk = wave_sum(acc.kills);
  double v = acc.maxdd;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
  const bool l0 = threadIdx.x == 0;
  unsigned long long* c0;
  unsigned long long* c1;
  unsigned long long* c2;
  unsigned long long* c3;
  if (GEN) {
    c0 = &g.state->pair_tests;
    c1 = &g.state->rechecks;
    c2 = mode == PM_GEN_OLD ? &g.state->rej_old : &g.state->rej_new;
    c3 = (unsigned long long*)&g.state->max_dd;
  } else {
    c0 = &a.ctr[0];
    c1 = &a.ctr[1];
    c2 = &a.ctr[3];  // unused for contacts
    c3 = &a.ctr[2];
  }
  // every lane issues the atomics (lane 0 carries the value) so no branch
  // precedes the next cross-lane operation of the caller
  atomicAdd(c0, (unsigned long long)(l0 ? wt : 0u));
  atomicAdd(c1, (unsigned long long)(l0 ? wr : 0u));
  if (GEN) atomicAdd(c2, (unsigned long long)(l0 ? wk : 0u));
  atomicMax(c3, (unsigned long long)(l0 ? __double_as_longlong(v) : 0ll));
  unsigned long long* cb =
      GEN ? (mode == PM_GEN_OLD ? &g.state->tile_bytes : &g.state->tile_bytes_new) : &a.ctr[3];
  atomicAdd(cb, l0 ? acc.bytes : 0ull);
  acc.tests = acc.rech = acc.kills = 0;
  acc.maxdd = 0.0;
  acc.bytes = 0;
}

template <int MODE, int PREC>
__device__ void process_cell(const PairArgs& a, const CellGeom& g, uint32_t gc, float4* tP,
                             float4* tD, uint32_t* tX, uint32_t* tDummy, Acc& acc) {
  constexpr bool GEN = (MODE == PM_GEN_OLD || MODE == PM_GEN_NEW);
  const int lane = threadIdx.x;
  const uint32_t pc = uni(a.prCount[gc]);
  if (pc == 0) return;
  const uint32_t ps = uni(a.prStart[gc]);
  const uint32_t lc = gc - g.cell_base;
  const int c0 = (int)(lc % (uint32_t)g.nc[0]);
  const int c1 = (int)((lc / (uint32_t)g.nc[0]) % (uint32_t)g.nc[1]);
  const int c2 = (int)(lc / ((uint32_t)g.nc[0] * (uint32_t)g.nc[1]));

  for (uint32_t pb = 0; pb < pc; pb += WAVE) {
    const uint32_t G = min((uint32_t)WAVE, pc - pb);
    uint32_t Gp = 1;
    while (Gp < G) Gp <<= 1;
    const uint32_t K = WAVE / Gp;
    const uint32_t lk = __builtin_ctz(K);
    const uint32_t p = (uint32_t)lane >> lk;
    const uint32_t sub = (uint32_t)lane & (K - 1);
    const bool inrange = p < G;
    const uint32_t pidx = ps + pb + (inrange ? p : 0u);  // clamped: every lane loads
    const float4 P1 = a.prP[pidx];
    const float4 D1 = a.prD[pidx];
    const uint32_t X1 = a.prX[pidx];
    bool have = inrange;
    if (MODE == PM_GEN_NEW) have = have && (a.prAlive[pidx] != 0u);
    const float chunk_hl = unif(wave_max(have ? P1.w : 0.0f));
    acc.bytes += (unsigned long long)G * 36ull;
    bool hit = false;
    const float rr_chunk = (chunk_hl + g.hl_max) * 1.000001f + g.R_max + CELL_MARGIN;
    const float rr_chunk2 = rr_chunk * rr_chunk;

    for (uint32_t si = 0; si < g.st_n; ++si) {
      const int4 so = g.st[si];
      const int ox = unii(so.x), oy = unii(so.y), oz = unii(so.z);
      const float md2 = unif(__int_as_float(so.w));
      if (md2 > rr_chunk2) break;
      const int raw[3] = {c0 + ox, c1 + oy, c2 + oz};
      int img[3], m[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        int mm = raw[d] % g.nc[d];
        if (mm < 0) mm += g.nc[d];
        img[d] = (raw[d] - mm) / g.nc[d];
        m[d] = mm;
      }
      const uint32_t nc = g.cell_base + (uint32_t)(m[0] + g.nc[0] * (m[1] + g.nc[1] * m[2]));
      uint32_t n = uni(a.nbCount[nc]);
      if (MODE == PM_GEN_OLD) n = min(n, g.cap);
      if (n == 0) continue;
      const float rr_c = (chunk_hl + unif(a.nbHl[nc])) * 1.000001f + g.R_max + CELL_MARGIN;
      if (md2 > rr_c * rr_c) continue;
      const double sx = (double)img[0] * g.box[0], sy = (double)img[1] * g.box[1],
                   sz = (double)img[2] * g.box[2];
      const float fx = (float)sx, fy = (float)sy, fz = (float)sz;
      const bool same = (ox == 0 && oy == 0 && oz == 0);
      const uint32_t nstart = uni(a.nbStart[nc]);
      for (uint32_t tb = 0; tb < n; tb += WAVE) {
        const uint32_t tn = min((uint32_t)WAVE, n - tb);
        acc.bytes += (unsigned long long)tn * 36ull;
        // every lane loads (index clamped to the tile) and writes its slot;
        // slots >= tn are never read
        const uint32_t q = nstart + tb + min((uint32_t)lane, tn - 1u);
        const float4 lp = a.nbP[q];
        const float4 ld = a.nbD[q];
        uint32_t x = a.nbX[q];
        if (MODE == PM_GEN_NEW) x = (a.nbAlive[q] != 0u) ? x : 0xFFFFFFFFu;
        tP[lane] = lp;
        tD[lane] = ld;
        tX[lane] = x;
        __syncthreads();
        if (have) {
          for (uint32_t j = sub; j < tn; j += K) {
            const uint32_t xj = tX[j];
            if (MODE == PM_GEN_NEW && xj >= X1) continue;
            if (MODE == PM_CNT_COUNT || MODE == PM_CNT_RECORD) {
              if (same && tb + j <= pb + p) continue;
              if (xj == X1) continue;
            }
            const float4 P2 = tP[j], D2 = tD[j];
            // FP64 evaluation in the canonical orientation the decisions are
            // defined by: generation puts the probe (proposal, or the later
            // slot) first; contacts put the lower rod id first, with the
            // periodic image applied to the higher id.  The closest-point
            // solve is not bitwise symmetric under swapping the segments.
            const bool swap = !GEN && (X1 > xj);
            const float4 Pa = swap ? P2 : P1, Da = swap ? D2 : D1;
            const float4 Pb = swap ? P1 : P2, Db = swap ? D1 : D2;
            const double ux = swap ? -sx : sx, uy = swap ? -sy : sy, uz = swap ? -sz : sz;
            bool within;
            bool cut;
            double d2_64 = -1.0;
            double R64;
            if (GEN)
              R64 = g.sf64 * ((double)D1.w + (double)D2.w) + g.sep64;
            else
              R64 = ((double)Da.w + (double)Db.w) + a.gap64;
            if (PREC == NF_PREC_FP64) {
              d2_64 = pair_d2_f64(Pa, Da, Pb, Db, ux, uy, uz, R64, &within);
              cut = GEN ? (d2_64 < R64 * R64) : (d2_64 <= R64 * R64);
            } else {
              const float R32 = GEN ? (g.sf32 * (D1.w + D2.w) + g.sep32) : ((D1.w + D2.w) + a.gap32);
              const F32Pair r = pair_f32(P1, D1, P2, D2, fx, fy, fz, R32, g.ew, GEN);
              if (r.decided < 0) {
                ++acc.rech;
                d2_64 = pair_d2_f64(Pa, Da, Pb, Db, ux, uy, uz, R64, &within);
                cut = GEN ? (d2_64 < R64 * R64) : (d2_64 <= R64 * R64);
              } else {
                within = r.within != 0;
                cut = r.decided != 0;
                if (a.diag && within) {
                  bool w2;
                  const double e64 = pair_d2_f64(Pa, Da, Pb, Db, ux, uy, uz, R64, &w2);
                  const double dd = fabs(sqrt((double)r.d2) - sqrt(e64));
                  acc.maxdd = dd > acc.maxdd ? dd : acc.maxdd;
                }
              }
            }
            if (!within) continue;
            ++acc.tests;
            if (!cut) continue;
            if (GEN) {
              hit = true;
            } else {
              if (d2_64 < 0.0) {
                bool w2;
                d2_64 = pair_d2_f64(Pa, Da, Pb, Db, ux, uy, uz, R64, &w2);
              }
              const uint32_t I = X1 < xj ? X1 : xj;
              const uint32_t J = X1 < xj ? xj : X1;
              if (MODE == PM_CNT_COUNT) {
                atomicAdd(&a.cnt[I], 1u);
              } else {
                const int sgn = X1 < xj ? 1 : -1;
                const uint32_t k = atomicAdd(&a.fill[I], 1u);
                if (k >= a.rowptr[I + 1] - a.rowptr[I]) continue;  // host reports the mismatch
                const uint32_t pos = a.rowptr[I] + k;
                const double gv = sqrt_cr(d2_64) - ((double)Da.w + (double)Db.w);
                a.outJ[pos] = J;
                a.outGap[pos] = (float)gv;
                a.outGap64[pos] = gv;
                a.outImg[pos] = pack_img(sgn * img[0], sgn * img[1], sgn * img[2]);
              }
            }
          }
        }
        __syncthreads();
      }
    }
    if (GEN) {
      const unsigned long long mk = __ballot(hit);
      const unsigned long long gm = (K == 64) ? ~0ull : ((1ull << K) - 1ull);
      const bool hitp = ((mk >> (p * K)) & gm) != 0ull;
      const bool lead = (sub == 0) && have;
      acc.kills += (lead && hitp) ? 1u : 0u;
      // branch-free result store: lanes without a result write a private
      // LDS slot instead
      if (MODE == PM_GEN_OLD) {
        uint32_t* dst = (lead && hitp) ? &a.alive_out[pidx] : &tDummy[lane];
        *dst = 0u;
      } else {
        uint32_t* dst = (have && !hitp) ? &a.surv[X1] : &tDummy[lane];
        *dst = 1u;
      }
    }
  }
}

template <int MODE, int PREC>
__global__ void __launch_bounds__(PAIR_THREADS) k_pair(PairArgs a) {
  __shared__ float4 tP[WAVE];
  __shared__ float4 tD[WAVE];
  __shared__ uint32_t tX[WAVE];
  __shared__ uint32_t tDummy[WAVE];
  constexpr bool GEN = (MODE == PM_GEN_OLD || MODE == PM_GEN_NEW);
  const uint32_t W = a.ncells_total;
  const int xcd = blockIdx.x & 7;
  // rounds after every instance finished plan no candidates: nothing to do
  if (GEN && uni(a.ri->T) == 0u) return;
  const bool l0 = threadIdx.x == 0;
  CellGeom g;
  Acc acc = {0u, 0u, 0u, 0.0, 0ull};
  uint32_t cur_inst = 0xFFFFFFFFu;
  if (!GEN) {
    for (int d = 0; d < 3; ++d) {
      g.box[d] = a.box[d];
      g.boxf[d] = a.boxf[d];
      g.nc[d] = a.ncell[d];
    }
    g.cell_base = 0;
    g.st = a.stencil + a.st_off;
    g.st_n = a.st_n;
    g.ew = a.ew;
    g.hl_max = a.hl_max;
    g.R_max = a.R_max;
    g.sf64 = 1.0;
    g.sep64 = 0.0;
    g.sf32 = 1.0f;
    g.sep32 = 0.0f;
    g.cap = 0xFFFFFFFFu;
    g.state = nullptr;
  }
  for (int k = 0; k < 8; ++k) {
    const int q = (xcd + k) & 7;
    const uint32_t qb = (uint32_t)(((uint64_t)W * (uint32_t)q) / 8u);
    const uint32_t qe = (uint32_t)(((uint64_t)W * (uint32_t)(q + 1)) / 8u);
    // a plain read first: an exhausted queue costs no atomic
    if (uni(__hip_atomic_load(&a.qctr[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >= qe - qb)
      continue;
    for (;;) {
      // all lanes take part in the atomic (lane 0 adds 1), lane 0's old
      // value is the work item
      const uint32_t old = atomicAdd(&a.qctr[q], l0 ? 1u : 0u);
      const uint32_t item = uni(old);
      if (item >= qe - qb) break;
      const uint32_t gc = qb + item;
      if (GEN) {
        const uint32_t inst = uni(a.cell_inst[gc]);
        if (inst != cur_inst) {
          if (cur_inst != 0xFFFFFFFFu) flush<GEN>(a, g, acc, MODE);
          const InstDesc& d = a.desc[inst];
          for (int dd = 0; dd < 3; ++dd) {
            g.box[dd] = d.smp.box[dd];
            g.boxf[dd] = d.boxf[dd];
            g.nc[dd] = d.ncell[dd];
          }
          g.cell_base = d.cell_base;
          g.st = a.stencil + d.st_off;
          g.st_n = d.st_n;
          g.ew = d.ew;
          g.hl_max = d.hl_max;
          g.R_max = d.R_max;
          g.sf64 = d.sf64;
          g.sep64 = d.sep64;
          g.sf32 = d.sf32;
          g.sep32 = d.sep32;
          g.cap = d.cap;
          g.state = &a.st[inst];
          cur_inst = inst;
        }
      }
      process_cell<MODE, PREC>(a, g, gc, tP, tD, tX, tDummy, acc);
    }
  }
  if (!GEN || cur_inst != 0xFFFFFFFFu) flush<GEN>(a, g, acc, MODE);
}

static int g_pair_grid = 0;

int pair_grid() {
  if (g_pair_grid == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    g_pair_grid = cus * 16;
  }
  return g_pair_grid;
}

void launch_pair(int mode, int prec, const PairArgs& a, hipStream_t s) {
  const dim3 grid(pair_grid()), block(PAIR_THREADS);
#define NF_LAUNCH(M, P) hipLaunchKernelGGL((k_pair<M, P>), grid, block, 0, s, a)
  if (prec == NF_PREC_FP64) {
    switch (mode) {
      case PM_GEN_OLD: NF_LAUNCH(PM_GEN_OLD, NF_PREC_FP64); break;
      case PM_GEN_NEW: NF_LAUNCH(PM_GEN_NEW, NF_PREC_FP64); break;
      case PM_CNT_COUNT: NF_LAUNCH(PM_CNT_COUNT, NF_PREC_FP64); break;
      default: NF_LAUNCH(PM_CNT_RECORD, NF_PREC_FP64); break;
    }
  } else {
    switch (mode) {
      case PM_GEN_OLD: NF_LAUNCH(PM_GEN_OLD, NF_PREC_FP32_RECHECK); break;
      case PM_GEN_NEW: NF_LAUNCH(PM_GEN_NEW, NF_PREC_FP32_RECHECK); break;
      case PM_CNT_COUNT: NF_LAUNCH(PM_CNT_COUNT, NF_PREC_FP32_RECHECK); break;
      default: NF_LAUNCH(PM_CNT_RECORD, NF_PREC_FP32_RECHECK); break;
    }
  }
#undef NF_LAUNCH
}

}  // namespace nf
