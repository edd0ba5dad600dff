// nf_seg.hip - the footprint contact pass (nf_seg.h) over the centre cell
// list: one wave per work unit (a run of home cells along x).
//
// 0. Work: a unit is a run of SEG_ROW home cells along x.  Units are taken
//    from eight counters, one per XCD and each on its own 128-byte line;
//    counter q covers the q-th of eight equal ranges of the unit index (z
//    slabs, so a wave's neighbours stay in its XCD's L2), and a wave turns
//    to the next XCD's counter when its own range is used up.  One counter
//    for all cells serialised every grab on one line.
// 1. Gather, once per unit: walk the run's stencil (sorted by minimum
//    distance, cells within hl_max + r) 64 offsets at a time; the rods of
//    the block's relevant cells are packed back to back and staged 64 at a
//    time (a lane finds its cell from scalar splitters and one read of
//    eight of the block's prefix counts in LDS), so a batch is full however
//    few rods a cell holds, and the next batch's loads are in flight while
//    one is tested.  Rods whose
//    capsule (radius r) crosses one of the run's cells are the unit's
//    survivors (at most SMAX, kept in LDS with their periodic image and the
//    cells they cross).  The test reads each rod's near mask (k_seg_marks,
//    seg_near_mask: the cells of the 3 x 3 x 3 block around its own cell
//    that it crosses, computed once per rod) and the listed cell's offset
//    from the run (seg_near_cells): 4 bytes and a few bit operations per
//    examined rod.  Only a rod that may leave its block (SEG_LONG) loads its
//    axis for the crossing test (seg_row_cells); the survivors' rods are
//    loaded after the walk.  A run shares one gather among its cells:
//    about 4x fewer rods examined than a gather per cell.
// Steps 2 and 3 run per home cell of the unit, on its survivors.
// 2. Fine masks: each survivor's footprint on the f x f x f fine cells of the
//    home cell (seg_fine_mask), and per fine cell a bit set of its survivors.
// 3. Pairs: survivor i's candidates are the later survivors sharing one of
//    its fine cells (the union of those cells' bit sets, so a pair meeting
//    in several fine cells is listed once); they queue in LDS and run 64 at a
//    time through the FP32 screen (seg_screen).  The few it leaves open go
//    to a second LDS queue and run 64 at a time through the band, FP64 and
//    the reporting rule (seg_full), which keeps a contact only in the home
//    cell that holds the midpoint of its closest points.  Every contact is
//    therefore found (its midpoint lies within r of both axes, so both rods
//    survive in that cell and share the midpoint's fine cell) and reported
//    once.
//
// A unit with more than SMAX survivors is redone one home cell at a time; a
// single cell with more raises the overflow flag and is skipped, and the
// host then runs the cell-pair engine instead.  Control flow
// outside seg_screen and seg_full is wave-uniform; cross-lane operations run
// with every lane active.
#include "nf_internal.h"

namespace nf {

namespace {

constexpr int SMAX = 256;       // survivors per work unit
constexpr int SWORDS = SMAX / 64;
constexpr int FMAX = 64;        // fine cells per home cell (f <= 4)
constexpr int XQ = 2 * WAVE;    // candidate pairs (fewer than 64 wait between pushes of up to 64)
constexpr uint32_t XQUEUES = 8; // one work counter per XCD

struct XLds {
  float4 sP[SMAX], sD[SMAX];
  uint32_t sX[SMAX], sM[SMAX];  // rod id, image (signed bytes)
  uint8_t sC[SMAX];             // home cells of the unit a survivor crosses (bit i: the unit's cell i)
  uint8_t lst[SMAX];            // the survivors of the current home cell
  unsigned long long fm[SMAX];  // fine masks of the current cell's survivors, by list position
  unsigned long long mem[FMAX][SWORDS];
  // the gather's block tables and the pair queues are never live together
  // (barriers between the phases): one shares the other's storage, so that
  // ten waves fit a CU's LDS
  union {
    struct {
      uint32_t q[XQ];   // candidate pairs: survivor i | j << 8
      uint32_t q2[XQ];  // screened pairs: i | j << 8 | seg_screen code << 16
    };
    struct {
      // gather: prefix counts, starts, images, offsets from the run's first cell
      alignas(16) uint32_t gpre[WAVE];  // read eight at a time from a multiple of eight
      uint32_t gst[WAVE], gmeta[WAVE], gwin[WAVE];
    };
  };
};

__device__ __forceinline__ uint32_t xuni(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ uint32_t xprefix(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ int ximg(uint32_t meta, int a) { return (int)(int8_t)((meta >> (8 * a)) & 0xFFu); }

// Second stage: the pair of screened entry e (valid lanes only) through the
// band or FP64 and the reporting rule; appends a contact.  No cross-lane
// operation inside.
template <int PREC>
__device__ __forceinline__ void x_full(const SegCrossArgs& a, XLds& L, uint32_t e, bool valid, uint32_t cell,
                                       uint32_t& tests, uint32_t& rech, double& maxdd) {
  const uint32_t ent = L.q2[e];
  const uint32_t i = ent & 0xFFu, j = (ent >> 8) & 0xFFu;
  const int code = (int)(ent >> 16);
  const float4 Pi = L.sP[i], Di = L.sD[i], Pj = L.sP[j], Dj = L.sD[j];
  const uint32_t xi = L.sX[i], xj = L.sX[j], mi = L.sM[i], mj = L.sM[j];
  if (!valid) return;
  SegHit hit;
  if (!seg_full<PREC>(a.x, a.g, xi, Pi, Di, mi, xj, Pj, Dj, mj, cell, code, tests, rech, maxdd, &hit)) return;
  const uint32_t pos = atomicAdd(a.app_n, 1u);
  if (pos >= a.app_cap) return;
  a.outI[pos] = hit.I;
  a.outJ[pos] = hit.J;
  a.outGap[pos] = (float)hit.gap;
  a.outGap64[pos] = hit.gap;
  a.outImg[pos] = hit.img;
}

// First stage: candidate entries [base, base + cnt) through the FP32
// screen; those it leaves open go to the second queue, which runs when 64
// wait.  Wave-uniform control; the caller holds a barrier before the call
// and one after it.
template <int PREC>
__device__ __forceinline__ void x_screen(const SegCrossArgs& a, XLds& L, uint32_t base, uint32_t cnt,
                                         uint32_t& q2n, uint32_t cell, uint32_t& tests, uint32_t& rech,
                                         double& maxdd) {
  const uint32_t lane = threadIdx.x;
  const bool valid = lane < cnt;
  const uint32_t ent = L.q[base + (valid ? lane : 0u)];
  const uint32_t i = ent & 0xFFu, j = (ent >> 8) & 0xFFu;
  const float4 Pi = L.sP[i], Di = L.sD[i], Pj = L.sP[j], Dj = L.sD[j];
  const uint32_t xi = L.sX[i], xj = L.sX[j], mi = L.sM[i], mj = L.sM[j];
  // a rod and its own periodic image are not a pair (contacts have I < J)
  const int code = (valid && xi != xj) ? seg_screen<PREC>(a.x, Pi, Di, mi, Pj, Dj, mj, tests) : 0;
  const bool f = code != 0;
  const unsigned long long m = __ballot(f);
  if (f) L.q2[q2n + xprefix(m)] = (ent & 0xFFFFu) | ((uint32_t)code << 16);
  q2n += (uint32_t)__popcll(m);
  if (q2n >= (uint32_t)WAVE) {
    __syncthreads();
    x_full<PREC>(a, L, q2n - WAVE + lane, true, cell, tests, rech, maxdd);
    q2n -= WAVE;
  }
}

// Gather of the run of ws home cells from x = xs in row (h1, h2): the rods
// whose capsule crosses one of them go to the survivor table with the cells
// they cross.  Returns their count; past SMAX only the first SMAX are kept.
__device__ __forceinline__ uint32_t x_gather(const SegCrossArgs& a, XLds& L, int xs, int ws, int h1, int h2,
                                             float mx, unsigned long long& bytes) {
  const uint32_t lane = threadIdx.x;
  float lo[3], hi[3];
  lo[0] = (float)xs * a.g.h[0];
  hi[0] = (float)(xs + ws) * a.g.h[0];
  lo[1] = (float)h1 * a.g.h[1];
  hi[1] = (float)(h1 + 1) * a.g.h[1];
  lo[2] = (float)h2 * a.g.h[2];
  hi[2] = (float)(h2 + 1) * a.g.h[2];
  uint32_t S = 0;
  for (uint32_t sb = 0; sb < a.st_n; sb += WAVE) {
    const uint32_t si = sb + lane;
    const bool sv = si < a.st_n;
    const int4 so = a.stencil[sv ? si : a.st_n - 1u];
    const float md2 = __int_as_float(so.w);
    const int raw[3] = {xs + so.x, h1 + so.y, h2 + so.z};
    int m[3], im[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int q = floor_div(raw[k], a.g.nc[k]);
      im[k] = q;
      m[k] = raw[k] - q * a.g.nc[k];
    }
    const uint32_t nc = (uint32_t)(m[0] + a.g.nc[0] * (m[1] + a.g.nc[1] * m[2]));
    const uint32_t n = a.ccount[nc];
    const float hl = a.chl[nc];
    const uint32_t st0 = a.cstart[nc];
    // the three loads in flight together (the compiler would otherwise load
    // the longest rod only behind a branch on the count: a second round trip)
    asm volatile("" ::"v"(n), "v"(hl), "v"(st0));
    const float rr = hl * 1.000001f + a.rpad;
    const bool rel = sv && n > 0u && md2 <= rr * rr;
    const bool beyond = sv && md2 > a.reach2;
    const uint32_t meta = seg_pack(im[0], im[1], im[2]);
    // the listed cell's offset from the run's first cell, clamped to what a
    // near mask can tell apart (a rod crosses only cells within one of its own)
    const uint32_t win = seg_pack(max(-2, min(so.x, SEG_ROW + 1)), max(-2, min(so.y, 2)), max(-2, min(so.z, 2)));
    // the block's relevant cells back to back: inclusive prefix of counts
    const uint32_t cnt = rel ? n : 0u;
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)incl, o, WAVE);
      incl += lane >= (uint32_t)o ? t : 0u;
    }
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, WAVE - 1);
    const uint32_t pre = incl - cnt;
    L.gpre[lane] = pre;
    L.gst[lane] = st0;
    L.gmeta[lane] = meta;
    L.gwin[lane] = win;
    // every eighth prefix in scalar registers: the first level of the search
    uint32_t spl[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) spl[m] = (uint32_t)__builtin_amdgcn_readlane((int)pre, 8 * (m + 1));
    __syncthreads();
    // rod pos of the packed list: the last lane whose exclusive prefix is
    // <= pos holds it (lanes without rods share the next lane's prefix; the
    // prefixes do not decrease).  Its group of eight lanes from the scalar
    // splitters, then its place in the group from one read of the group's
    // eight prefixes: two LDS round trips instead of a six-step search.
    // An empty cell's start can be one past the last rod, so idle lanes
    // load rod 0's mask, which exists whenever total > 0.
    uint32_t src = 0u, mk = 0u, cm = 0u, wn = 0u;
    bool v = false;
    auto fetch = [&](uint32_t b0) {
      const uint32_t pos = b0 + lane;
      v = pos < total;
      uint32_t k = 0;
#pragma unroll
      for (int m = 0; m < 7; ++m) k += spl[m] <= pos ? 8u : 0u;
      const uint4 g0 = *reinterpret_cast<const uint4*>(&L.gpre[k]);
      const uint4 g1 = *reinterpret_cast<const uint4*>(&L.gpre[k + 4u]);
      k += (uint32_t)(g0.y <= pos) + (uint32_t)(g0.z <= pos) + (uint32_t)(g0.w <= pos) + (uint32_t)(g1.x <= pos) +
           (uint32_t)(g1.y <= pos) + (uint32_t)(g1.z <= pos) + (uint32_t)(g1.w <= pos);
      cm = L.gmeta[k];
      wn = L.gwin[k];
      src = v ? L.gst[k] + (pos - L.gpre[k]) : 0u;
      mk = a.cmask[src];
    };
    if (total > 0u) fetch(0u);
    for (uint32_t b0 = 0; b0 < total; b0 += WAVE) {
      // this batch, then the next batch's loads in flight while it is tested
      const uint32_t bsrc = src, bmk = mk, bcm = cm, bwn = wn;
      const bool bv = v;
      if (b0 + WAVE < total) fetch(b0 + WAVE);
      uint32_t cells = 0u;
      if (bv) {
        bytes += 4ull;
        if ((bmk & SEG_LONG) == 0u) {
          cells = seg_near_cells(bmk, ximg(bwn, 0), ximg(bwn, 1), ximg(bwn, 2), ws);
        } else {
          // a rod that may reach beyond its neighbour cells: the crossing test
          const float4 bP = a.cP[bsrc], bD = a.cD[bsrc];
          bytes += 32ull;
          const float p[3] = {bP.x + (float)ximg(bcm, 0) * a.x.boxf[0], bP.y + (float)ximg(bcm, 1) * a.x.boxf[1],
                              bP.z + (float)ximg(bcm, 2) * a.x.boxf[2]};
          const float c[3] = {bD.x, bD.y, bD.z};
          cells = seg_row_cells(p, c, bP.w, lo, hi, xs, ws, a.g.h[0], a.g.r, mx);
        }
      }
      const bool keep = cells != 0u;
      const unsigned long long b = __ballot(keep);
      const uint32_t at = S + xprefix(b);
      if (keep && at < (uint32_t)SMAX) {
        L.sX[at] = bsrc;  // list position until the survivors are loaded
        L.sM[at] = bcm;
        L.sC[at] = (uint8_t)cells;
      }
      S += (uint32_t)__popcll(b);
    }
    __syncthreads();
    if (__ballot(beyond) != 0ull) break;
  }
  // the survivors' rods, from their list positions
  if (S <= (uint32_t)SMAX) {
    for (uint32_t k = lane; k < S; k += WAVE) {
      const uint32_t at = L.sX[k];
      L.sP[k] = a.cP[at];
      L.sD[k] = a.cD[at];
      L.sX[k] = a.cid[at];
    }
    bytes += 36ull * (uint64_t)((S > lane) ? (S - lane + WAVE - 1u) / WAVE : 0u);
    __syncthreads();
  }
  return S;
}

// Steps 2 and 3 for each home cell of the gathered run (S <= SMAX
// survivors in the table).
template <int PREC>
__device__ __forceinline__ void x_cells(const SegCrossArgs& a, XLds& L, uint32_t S, int xs, int ws, int h1,
                                        int h2, const float e[3], const float ie[3], uint32_t& tests,
                                        uint32_t& rech, double& maxdd) {
  const uint32_t lane = threadIdx.x;
  const int f = a.fine;
  const uint32_t F = (uint32_t)(f * f * f);  // fine cells per home cell, at most FMAX
  for (int ci = 0; ci < ws; ++ci) {
    const uint32_t cell =
        (uint32_t)(xs + ci) + (uint32_t)a.g.nc[0] * ((uint32_t)h1 + (uint32_t)a.g.nc[1] * (uint32_t)h2);
    const float clo[3] = {(float)(xs + ci) * a.g.h[0], (float)h1 * a.g.h[1], (float)h2 * a.g.h[2]};
    // ---- 2. this cell's survivors, their fine masks and the fine cells' sets
    uint32_t Sc = 0;
    for (uint32_t w0 = 0; w0 < S; w0 += WAVE) {
      const uint32_t idx = w0 + lane;
      const bool in = idx < S && ((L.sC[idx] >> ci) & 1u) != 0u;
      const unsigned long long b = __ballot(in);
      if (in) L.lst[Sc + xprefix(b)] = (uint8_t)idx;
      Sc += (uint32_t)__popcll(b);
    }
    // the fine cells' sets start empty (F <= 64 rows, one per lane)
    if (lane < F) {
#pragma unroll
      for (int k = 0; k < SWORDS; ++k) L.mem[lane][k] = 0ull;
    }
    __syncthreads();
    for (uint32_t wd = 0; wd * WAVE < Sc; ++wd) {
      const uint32_t k = wd * WAVE + lane;
      if (k < Sc) {
        const uint32_t id = L.lst[k];
        const float4 P = L.sP[id], D = L.sD[id];
        const uint32_t cm = L.sM[id];
        const float p[3] = {P.x + (float)ximg(cm, 0) * a.x.boxf[0], P.y + (float)ximg(cm, 1) * a.x.boxf[1],
                            P.z + (float)ximg(cm, 2) * a.x.boxf[2]};
        const float c[3] = {D.x, D.y, D.z};
        const unsigned long long fm = seg_fine_mask(p, c, P.w, clo, e, ie, f, a.g.r);
        L.fm[k] = fm;
        // this survivor joins the set of each fine cell it touches (a few
        // LDS atomics per lane instead of a ballot per fine cell)
        unsigned long long t = fm;
        while (t != 0ull) {
          const int fc = (int)__builtin_ctzll(t);
          t &= t - 1ull;
          atomicOr(&L.mem[fc][wd], 1ull << lane);
        }
      }
    }
    __syncthreads();
    // ---- 3. candidate pairs through the two queues
    const uint32_t nw = (Sc + WAVE - 1) / WAVE;
    uint32_t qn = 0, q2n = 0;
    for (uint32_t wd = 0; wd < nw; ++wd) {
      const uint32_t i = wd * WAVE + lane;
      const bool v = i < Sc;
      const uint32_t si = v ? (uint32_t)L.lst[i] : 0u;
      // partners: later list positions sharing a fine cell
      unsigned long long fm = v ? L.fm[i] : 0ull;
      unsigned long long part[SWORDS];
#pragma unroll
      for (int k = 0; k < SWORDS; ++k) part[k] = 0ull;
      while (fm != 0ull) {
        const int fc = (int)__builtin_ctzll(fm);
        fm &= fm - 1ull;
#pragma unroll
        for (int k = 0; k < SWORDS; ++k) part[k] |= L.mem[fc][k];
      }
      // keep j > i among the filled words
#pragma unroll
      for (int k = 0; k < SWORDS; ++k) {
        if ((uint32_t)k < wd || (uint32_t)k >= nw) part[k] = 0ull;
        if ((uint32_t)k == wd) part[k] &= (lane == 63u) ? 0ull : (~0ull << (lane + 1u));
      }
      for (;;) {
        // this lane's next partner: the lowest set bit of the first
        // non-empty word (unrolled: the words stay in registers)
        bool have = false;
        uint32_t j = 0;
#pragma unroll
        for (int k = 0; k < SWORDS; ++k) {
          if (!have && part[k] != 0ull) {
            j = (uint32_t)(k * 64) + (uint32_t)__builtin_ctzll(part[k]);
            part[k] &= part[k] - 1ull;
            have = true;
          }
        }
        const unsigned long long bm = __ballot(have);
        if (bm == 0ull) break;
        if (have) L.q[qn + xprefix(bm)] = si | ((uint32_t)L.lst[j] << 8);
        qn += (uint32_t)__popcll(bm);
        if (qn >= (uint32_t)WAVE) {
          __syncthreads();
          x_screen<PREC>(a, L, qn - WAVE, WAVE, q2n, cell, tests, rech, maxdd);
          qn -= WAVE;
          __syncthreads();
        }
      }
    }
    if (qn > 0) {
      __syncthreads();
      x_screen<PREC>(a, L, 0, qn, q2n, cell, tests, rech, maxdd);
      __syncthreads();
    }
    if (q2n > 0) x_full<PREC>(a, L, min(lane, q2n - 1u), lane < q2n, cell, tests, rech, maxdd);
    __syncthreads();
  }
}

template <int PREC>
__global__ void __launch_bounds__(WAVE) k_seg_cross(SegCrossArgs a) {
  __shared__ XLds L;
  const uint32_t lane = threadIdx.x;
  float e[3], ie[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    e[k] = a.g.h[k] / (float)a.fine;
    ie[k] = 1.0f / e[k];
  }
  const uint32_t nc0 = (uint32_t)a.g.nc[0], nc1 = (uint32_t)a.g.nc[1];
  const uint32_t nseg = (nc0 + SEG_ROW - 1u) / SEG_ROW;
  const uint32_t nunits = nseg * nc1 * (uint32_t)a.g.nc[2];
  const float mx = 4.0f * a.x.ew;  // rounding of a survivor's x extent (seg_row_cells)
  uint32_t tests = 0, rech = 0;
  double maxdd = 0.0;
  unsigned long long bytes = 0;
  const uint32_t xcd = blockIdx.x & (XQUEUES - 1u);
  uint32_t qi = 0, c0 = 0, c1 = 0;
  for (;;) {
    while (c0 == c1 && qi < XQUEUES) {
      const uint32_t q = (xcd + qi) & (XQUEUES - 1u);
      const uint32_t lo = (uint32_t)((unsigned long long)nunits * q / XQUEUES);
      const uint32_t hi = (uint32_t)((unsigned long long)nunits * (q + 1u) / XQUEUES);
      uint32_t old = 0u;
      if (lane == 0) old = atomicAdd(a.next + q * QSTRIDE, 1u);
      const uint32_t got = xuni(old);
      if (got < hi - lo) {
        c0 = lo + got;
        c1 = c0 + 1u;
      } else {
        ++qi;
      }
    }
    if (c0 == c1) break;
    // work unit: cells x0 .. x0 + w - 1 of one row
    const uint32_t u = c0++;
    const uint32_t row = u / nseg;
    const int x0 = (int)((u - row * nseg) * (uint32_t)SEG_ROW);
    const int w = min(SEG_ROW, (int)nc0 - x0);
    const int h1 = (int)(row % nc1), h2 = (int)(row / nc1);
    // the whole run, or, when its survivors overflow the table, one cell at
    // a time; a single cell that overflows raises the flag (the host then
    // runs the cell-pair engine)
    int xs = x0, ws = w;
    for (;;) {
      const uint32_t S = x_gather(a, L, xs, ws, h1, h2, mx, bytes);
      if (S <= (uint32_t)SMAX) {
        x_cells<PREC>(a, L, S, xs, ws, h1, h2, e, ie, tests, rech, maxdd);
      } else if (ws > 1) {
        if (lane == 0) atomicAdd(&a.ctr[5], 1ull);
        ws = 1;
        continue;
      } else {
        if (lane == 0) atomicOr(a.overflow, 1u);
        __syncthreads();
      }
      if (ws == 1 && xs + 1 < x0 + w) {
        ++xs;
        continue;
      }
      break;
    }
  }
  // one atomic per counter per wave, every lane converged
  const unsigned long long wt = wave_sum_u64(tests), wr = wave_sum_u64(rech), wb = wave_sum_u64(bytes);
  double v = maxdd;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
  if (lane == 0) {
    if (wt) atomicAdd(&a.ctr[0], wt);
    if (wr) atomicAdd(&a.ctr[1], wr);
    if (v > 0.0) atomicMax(&a.ctr[2], (unsigned long long)__double_as_longlong(v));
    atomicAdd(&a.ctr[3], wb);
  }
}

struct MarkArg {
  int32_t nc[3];
  double sc[3];
  float h[3];
  float r, mx;
};

// Near mask of every rod of the cell list (seg_near_mask), in list order.
__global__ void __launch_bounds__(256) k_seg_marks(const float4* __restrict__ cP, const float4* __restrict__ cD,
                                                   uint32_t n, MarkArg m, uint32_t* __restrict__ cmask) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 P = cP[i], D = cD[i];
    int cc[3];
    cell_coords(P, m.nc, m.sc, cc);
    const float p[3] = {P.x, P.y, P.z};
    const float c[3] = {D.x, D.y, D.z};
    cmask[i] = seg_near_mask(p, c, P.w, cc, m.h, m.r, m.mx);
  }
}

}  // namespace

void launch_seg_marks(const float4* cP, const float4* cD, uint32_t n, const int32_t nc[3], const double scale[3],
                      const float h[3], float r, float mx, uint32_t* cmask, hipStream_t s) {
  if (n == 0) return;
  MarkArg m;
  for (int k = 0; k < 3; ++k) {
    m.nc[k] = nc[k];
    m.sc[k] = scale[k];
    m.h[k] = h[k];
  }
  m.r = r;
  m.mx = mx;
  uint64_t blocks = ((uint64_t)n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_seg_marks, dim3((unsigned)blocks), dim3(256), 0, s, cP, cD, n, m, cmask);
}

// Persistent grid: 10 waves per CU, what the LDS holds (15.9 KB per wave;
// the registers would allow 12).
int seg_cross_grid() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  return cus * 10;
}

void launch_seg_cross(int prec, const SegCrossArgs& a, hipStream_t s) {
  const dim3 grid(seg_cross_grid()), block(WAVE);
  if (prec == NF_PREC_FP64)
    hipLaunchKernelGGL(k_seg_cross<NF_PREC_FP64>, grid, block, 0, s, a);
  else
    hipLaunchKernelGGL(k_seg_cross<NF_PREC_FP32_RECHECK>, grid, block, 0, s, a);
}

}  // namespace nf
