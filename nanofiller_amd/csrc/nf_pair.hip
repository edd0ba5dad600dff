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

template <int MODE, int PREC>
__device__ void process_cell(const PairArgs& a, const CellGeom& g, uint32_t gc, float4* tP,
                             float4* tD, uint32_t* tX) {
  constexpr bool GEN = (MODE == PM_GEN_OLD || MODE == PM_GEN_NEW);
  const int lane = threadIdx.x;
  const uint32_t pc = a.prCount[gc];
  if (pc == 0) return;
  const uint32_t ps = a.prStart[gc];
  const uint32_t lc = gc - g.cell_base;
  const int c0 = (int)(lc % (uint32_t)g.nc[0]);
  const int c1 = (int)((lc / (uint32_t)g.nc[0]) % (uint32_t)g.nc[1]);
  const int c2 = (int)(lc / ((uint32_t)g.nc[0] * (uint32_t)g.nc[1]));

  uint32_t tests = 0, rech = 0, kills = 0;
  double maxdd = 0.0;

  for (uint32_t pb = 0; pb < pc; pb += WAVE) {
    const uint32_t G = min((uint32_t)WAVE, pc - pb);
    uint32_t Gp = 1;
    while (Gp < G) Gp <<= 1;
    const uint32_t K = WAVE / Gp;
    const uint32_t lk = __builtin_ctz(K);
    const uint32_t p = (uint32_t)lane >> lk;
    const uint32_t sub = (uint32_t)lane & (K - 1);
    const uint32_t pidx = ps + pb + p;
    bool have = p < G;
    float4 P1 = make_float4(0.f, 0.f, 0.f, 0.f), D1 = P1;
    uint32_t X1 = 0;
    if (have) {
      P1 = a.prP[pidx];
      D1 = a.prD[pidx];
      X1 = a.prX[pidx];
      if (MODE == PM_GEN_NEW) have = a.prAlive[pidx] != 0u;
    }
    const float chunk_hl = wave_max(have ? P1.w : 0.0f);
    bool hit = false;
    const float rr_chunk = (chunk_hl + g.hl_max) * 1.000001f + g.R_max + CELL_MARGIN;
    const float rr_chunk2 = rr_chunk * rr_chunk;

    for (uint32_t si = 0; si < g.st_n; ++si) {
      const int4 so = g.st[si];
      const float md2 = __int_as_float(so.w);
      if (md2 > rr_chunk2) break;
      int raw[3] = {c0 + so.x, c1 + so.y, c2 + so.z};
      int img[3], m[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        int mm = raw[d] % g.nc[d];
        if (mm < 0) mm += g.nc[d];
        img[d] = (raw[d] - mm) / g.nc[d];
        m[d] = mm;
      }
      const uint32_t nc = g.cell_base + (uint32_t)(m[0] + g.nc[0] * (m[1] + g.nc[1] * m[2]));
      uint32_t n = a.nbCount[nc];
      if (MODE == PM_GEN_OLD) n = min(n, g.cap);
      if (n == 0) continue;
      const float rr_c = (chunk_hl + a.nbHl[nc]) * 1.000001f + g.R_max + CELL_MARGIN;
      if (md2 > rr_c * rr_c) continue;
      const double sx = (double)img[0] * g.box[0], sy = (double)img[1] * g.box[1],
                   sz = (double)img[2] * g.box[2];
      const float fx = (float)sx, fy = (float)sy, fz = (float)sz;
      const bool same = (so.x == 0 && so.y == 0 && so.z == 0);
      const uint32_t nstart = a.nbStart[nc];
      for (uint32_t tb = 0; tb < n; tb += WAVE) {
        const uint32_t tn = min((uint32_t)WAVE, n - tb);
        if ((uint32_t)lane < tn) {
          const uint32_t q = nstart + tb + lane;
          tP[lane] = a.nbP[q];
          tD[lane] = a.nbD[q];
          uint32_t x = a.nbX[q];
          if (MODE == PM_GEN_NEW && a.nbAlive[q] == 0u) x = 0xFFFFFFFFu;
          tX[lane] = x;
        }
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
            bool within;
            bool cut;
            double d2_64 = -1.0;
            double R64;
            if (GEN)
              R64 = g.sf64 * ((double)D1.w + (double)D2.w) + g.sep64;
            else
              R64 = ((double)D1.w + (double)D2.w) + a.gap64;
            if (PREC == NF_PREC_FP64) {
              d2_64 = pair_d2_f64(P1, D1, P2, D2, sx, sy, sz, R64, &within);
              cut = GEN ? (d2_64 < R64 * R64) : (d2_64 <= R64 * R64);
            } else {
              const float R32 = GEN ? (g.sf32 * (D1.w + D2.w) + g.sep32) : ((D1.w + D2.w) + a.gap32);
              const F32Pair r = pair_f32(P1, D1, P2, D2, fx, fy, fz, R32, g.ew, GEN);
              if (r.decided < 0) {
                ++rech;
                d2_64 = pair_d2_f64(P1, D1, P2, D2, sx, sy, sz, R64, &within);
                cut = GEN ? (d2_64 < R64 * R64) : (d2_64 <= R64 * R64);
              } else {
                within = r.within != 0;
                cut = r.decided != 0;
                if (a.diag && within) {
                  bool w2;
                  const double e64 = pair_d2_f64(P1, D1, P2, D2, sx, sy, sz, R64, &w2);
                  const double dd = fabs(sqrt((double)r.d2) - sqrt(e64));
                  if (dd > maxdd) maxdd = dd;
                }
              }
            }
            if (!within) continue;
            ++tests;
            if (!cut) continue;
            if (GEN) {
              hit = true;
            } else {
              if (d2_64 < 0.0) {
                bool w2;
                d2_64 = pair_d2_f64(P1, D1, P2, D2, sx, sy, sz, R64, &w2);
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
                const double gv = sqrt(d2_64) - ((double)D1.w + (double)D2.w);
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
      const unsigned long long m = __ballot(hit);
      const unsigned long long gm = (K == 64) ? ~0ull : ((1ull << K) - 1ull);
      const bool hitp = ((m >> (p * K)) & gm) != 0ull;
      const bool lead = (sub == 0) && have;
      if (lead && hitp) ++kills;
      if (MODE == PM_GEN_OLD && lead && hitp) a.alive_out[pidx] = 0u;
      if (MODE == PM_GEN_NEW && lead && !hitp) a.surv[X1] = 1u;
    }
  }
  const uint32_t wt = wave_sum(tests), wr = wave_sum(rech), wk = wave_sum(kills);
  if (a.diag) {
    double v = maxdd;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
    maxdd = v;
  }
  if (lane == 0) {
    if (GEN) {
      InstState* s = g.state;
      if (wt) atomicAdd(&s->pair_tests, (unsigned long long)wt);
      if (wr) atomicAdd(&s->rechecks, (unsigned long long)wr);
      if (wk) atomicAdd(MODE == PM_GEN_OLD ? &s->rej_old : &s->rej_new, (unsigned long long)wk);
      if (a.diag && maxdd > 0.0)
        atomicMax((unsigned long long*)&s->max_dd, (unsigned long long)__double_as_longlong(maxdd));
    } else {
      if (wt) atomicAdd(&a.ctr[0], (unsigned long long)wt);
      if (wr) atomicAdd(&a.ctr[1], (unsigned long long)wr);
      if (a.diag && maxdd > 0.0)
        atomicMax(&a.ctr[2], (unsigned long long)__double_as_longlong(maxdd));
    }
  }
}

template <int MODE, int PREC>
__global__ void __launch_bounds__(PAIR_THREADS) k_pair(PairArgs a) {
  __shared__ float4 tP[WAVE];
  __shared__ float4 tD[WAVE];
  __shared__ uint32_t tX[WAVE];
  constexpr bool GEN = (MODE == PM_GEN_OLD || MODE == PM_GEN_NEW);
  const uint32_t W = a.ncells_total;
  const int xcd = blockIdx.x & 7;
  CellGeom g;
  int cur_inst = -1;
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
    for (;;) {
      uint32_t item = 0;
      if (threadIdx.x == 0) item = atomicAdd(&a.qctr[q], 1u);
      item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, WAVE));
      if (item >= qe - qb) break;
      const uint32_t gc = qb + item;
      if (GEN) {
        const int inst = (int)a.cell_inst[gc];
        if (inst != cur_inst) {
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
      process_cell<MODE, PREC>(a, g, gc, tP, tD, tX);
    }
  }
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
