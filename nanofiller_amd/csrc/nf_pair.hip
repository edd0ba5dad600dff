// nf_pair.hip - the cell-pair engine.
//
// One 64-lane wavefront (one workgroup) per home cell, taken from 64 work
// queues (eight per XCD, each counter on its own cache line).  The home
// cell's probes are taken 64 at a time, one per lane, and read by the
// other lanes with readlane.  The stencil, sorted by minimum cell-to-cell
// distance, is looked up 64 offsets at a time (one round trip), and the
// neighbour cells that can reach the chunk are packed into a 128-slot
// staging buffer (two slots per lane).  Every (probe, slot) pair passes a
// conservative centre-distance prefilter; survivors queue in LDS, go
// through the rigorous FP32 screen 64 at a time, and those the screen
// cannot settle get the full FP32-band / FP64 test 64 at a time, so each
// expensive step runs on dense work.
//
// Modes replace the reference's kernels:
//   PM_GEN_OLD    new vs accepted rods   cuda_OldExtIsectLB (inc/intersection_kern.h:245-364)
//                 (run accepted-major, with the roles of the two sets
//                 swapped, once an instance's accepted rods outnumber the
//                 round's proposals four to one; otherwise a cut proposal
//                 leaves its stencil walk)
//   PM_GEN_NEW    new vs new, priority   cuda_InternalIsect_noLck / cuda_NewExtIsectLB_parted
//                                        (inc/intersection_kern.h:12-241)
//                 (over the first pass's survivors, compacted per cell;
//                 half stencil: each pair of them once; a cut clears the
//                 later slot's survivor flag)
//   PM_CNT        contact records        core_cell_interact_kernel CONTACT_COUNT and
//                                        CONTACT_RECORD in one pass (appended, unsorted)
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
  float rnc[3];  // 1 / nc (rounded): periodic wrap without integer division
  uint32_t cell_base;
  const int4* st;
  uint32_t st_n;
  float ew, hl_max, R_max;
  double sf64, sep64;
  float sf32, sep32;
  uint32_t cap;
  InstState* state;
  // home and neighbour rod sets of the pass; the reject-vs-accepted pass of
  // an instance whose accepted rods outnumber its proposals runs
  // accepted-major (inv): accepted rods are the probes, proposals the
  // staged neighbours, and a cut clears the proposal's alive flag directly
  const float4* hP;
  const float4* hD;
  const uint32_t* hX;
  const uint32_t* hStart;
  const uint32_t* hCount;
  const float4* nP;
  const float4* nD;
  const uint32_t* nX;
  const uint32_t* nStart;
  const uint32_t* nCount;
  const float* nHl;
  bool inv;
  bool idle;  // reject-vs-accepted pass: the instance has no proposals, or no accepted rods
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
    c2 = &g.state->rej_new;
    c3 = (unsigned long long*)&g.state->max_dd;
  } else {
    c0 = &a.ctr[0];
    c1 = &a.ctr[1];
    c2 = &a.ctr[3];  // unused for contacts
    c3 = &a.ctr[2];
  }
  // lane 0 alone issues the atomics: nothing in this branch is cross-lane,
  // and the caller's next cross-lane operation follows the reconvergence
  unsigned long long* cb =
      GEN ? (mode == PM_GEN_OLD ? &g.state->tile_bytes : &g.state->tile_bytes_new) : &a.ctr[3];
  if (l0) {
    atomicAdd(c0, (unsigned long long)wt);
    atomicAdd(c1, (unsigned long long)wr);
    // rejections by earlier proposals of the round (the round controller
    // counts those by accepted rods from the survivor counts)
    if (GEN && mode == PM_GEN_NEW) atomicAdd(c2, (unsigned long long)wk);
    atomicMax(c3, (unsigned long long)__double_as_longlong(v));
    atomicAdd(cb, acc.bytes);
  }
  acc.tests = acc.rech = acc.kills = 0;
  acc.maxdd = 0.0;
  acc.bytes = 0;
}

// LDS of one wave (under 10 KB, so 16 waves fit a CU): the probe chunk, the
// staged neighbour rods and the two pair queues.
constexpr int SLOTS = 2 * WAVE;
constexpr int QCAP = 3 * WAVE;
struct PairLds {
  float4 sP[WAVE], sD[WAVE];  // probes of the chunk
  uint32_t sX[WAVE];
  uint32_t hit[WAVE];          // per probe: overlaps something (generation)
  uint32_t dummy[WAVE];        // sink of branch-free stores
  float4 bP[SLOTS], bD[SLOTS];  // staged neighbour rods
  uint32_t bX[SLOTS];           // rod id (0xFFFFFFFF: not a candidate)
  uint32_t bM[SLOTS];           // image code | same cell << 31
  uint32_t qE[QCAP];            // queued pairs: probe | slot << 8
  uint32_t q2[QCAP];            // pairs for the full test: probe | slot << 8 | fp64 << 16
};

__device__ __forceinline__ uint32_t lane_prefix(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ int img_x(uint32_t meta) { return (int)(int8_t)(meta & 0xFF); }
__device__ __forceinline__ int img_y(uint32_t meta) { return (int)(int8_t)((meta >> 8) & 0xFF); }
__device__ __forceinline__ int img_z(uint32_t meta) { return (int)(int8_t)((meta >> 16) & 0xFF); }

// FP32 image shift of a neighbour: the FP64 shift img * box rounded once.
// Box edges are float-exact (nf_params_normalize) and |img| < 128, so the
// product is exact in FP64 and the FP32 product rounds the same exact value
// once: (float)img * boxf is that rounding, without FP64 arithmetic.
__device__ __forceinline__ float3 shift_f32(const CellGeom& g, uint32_t meta) {
  return make_float3((float)img_x(meta) * g.boxf[0], (float)img_y(meta) * g.boxf[1],
                     (float)img_z(meta) * g.boxf[2]);
}

// Screen of one queued pair (valid lanes only): the reach band and the
// line-distance bound of pair_f32_screen.  Counts the pair test when reach
// is decided.  Returns 0 (settled), 1 (full FP32 test needed) or 2 (reach
// undecided, or the plain FP64 path: the full test decides reach and cut in
// FP64 and counts).
template <int MODE, int PREC>
__device__ __forceinline__ uint32_t screen_test(const PairArgs& a, const CellGeom& g, PairLds& L,
                                                uint32_t e, bool valid, Acc& acc) {
  constexpr bool GEN = (MODE == PM_GEN_OLD || MODE == PM_GEN_NEW);
  const uint32_t ent = L.qE[e];
  const uint32_t p = ent & 0xFFu, j = ent >> 8;
  if (PREC == NF_PREC_FP64) return valid ? 2u : 0u;
  const float4 P2 = L.bP[j], D2 = L.bD[j];
  const uint32_t meta = L.bM[j];
  const float4 P1 = L.sP[p], D1 = L.sD[p];
  if (!valid) return 0u;
  const float3 sh = shift_f32(g, meta);
  const float R32 = GEN ? (g.sf32 * (D1.w + D2.w) + g.sep32) : ((D1.w + D2.w) + a.gap32);
  const int r = pair_f32_screen(P1, D1, P2, D2, sh.x, sh.y, sh.z, R32, g.ew);
  if (r == SCREEN_UNDECIDED) return 2u;
  if (r == SCREEN_OUT) return 0u;
  ++acc.tests;
  return r == SCREEN_FULL ? 1u : 0u;
}

// Full capsule test of one pair of the second queue (valid lanes only): the
// FP32 band with FP64 recheck, or plain FP64 (flag bit 16: reach not yet
// decided, count here).  No cross-lane operation inside.
template <int MODE, int PREC>
__device__ __forceinline__ void full_test(const PairArgs& a, const CellGeom& g, PairLds& L,
                                          uint32_t e, bool valid, Acc& acc) {
  constexpr bool GEN = (MODE == PM_GEN_OLD || MODE == PM_GEN_NEW);
  const uint32_t ent = L.q2[e];
  const uint32_t p = ent & 0xFFu, j = (ent >> 8) & 0xFFu;
  const bool f64 = PREC == NF_PREC_FP64 || ((ent >> 16) & 1u) != 0u;
  const float4 P2 = L.bP[j], D2 = L.bD[j];
  const uint32_t meta = L.bM[j];
  const uint32_t xj = L.bX[j];
  const float4 P1 = L.sP[p], D1 = L.sD[p];
  const uint32_t X1 = L.sX[p];
  if (!valid) return;
  const int ix = img_x(meta), iy = img_y(meta), iz = img_z(meta);
  const double sx = (double)ix * g.box[0], sy = (double)iy * g.box[1], sz = (double)iz * g.box[2];
  // FP64 evaluation in the canonical orientation the decisions are defined
  // by: generation puts the probe (proposal, or the later slot) first;
  // contacts put the lower rod id first, with the periodic image applied to
  // the higher id.  The closest-point solve is not bitwise symmetric under
  // swapping the segments.
  // generation: the later slot first (same-round pass), the proposal first
  // (reject-vs-accepted, swapped when run accepted-major)
  const bool swap = MODE == PM_GEN_NEW ? (X1 < xj) : (GEN ? g.inv : (X1 > xj));
  const float4 Pa = swap ? P2 : P1, Da = swap ? D2 : D1;
  const float4 Pb = swap ? P1 : P2, Db = swap ? D1 : D2;
  const double ux = swap ? -sx : sx, uy = swap ? -sy : sy, uz = swap ? -sz : sz;
  bool within = true;
  bool cut;
  double d2_64 = -1.0;
  double R64;
  if (GEN)
    R64 = g.sf64 * ((double)D1.w + (double)D2.w) + g.sep64;
  else
    R64 = ((double)Da.w + (double)Db.w) + a.gap64;
  if (f64) {
    if (PREC != NF_PREC_FP64) ++acc.rech;
    d2_64 = pair_d2_f64(Pa, Da, Pb, Db, ux, uy, uz, R64, &within);
    cut = GEN ? (d2_64 < R64 * R64) : (d2_64 <= R64 * R64);
    if (!within) return;
    ++acc.tests;
  } else {
    // reach was decided 'within' by the screen with the same band
    const float R32 = GEN ? (g.sf32 * (D1.w + D2.w) + g.sep32) : ((D1.w + D2.w) + a.gap32);
    const F32Pair r = pair_f32(P1, D1, P2, D2, (float)sx, (float)sy, (float)sz, R32, g.ew, GEN);
    if (r.decided < 0) {
      ++acc.rech;
      bool w64;
      d2_64 = pair_d2_f64(Pa, Da, Pb, Db, ux, uy, uz, R64, &w64);
      cut = GEN ? (d2_64 < R64 * R64) : (d2_64 <= R64 * R64);
    } else {
      cut = r.decided != 0;
      if (a.diag) {
        bool w2;
        const double e64 = pair_d2_f64(Pa, Da, Pb, Db, ux, uy, uz, R64, &w2);
        const double dd = fabs(sqrt((double)r.d2) - sqrt(e64));
        acc.maxdd = dd > acc.maxdd ? dd : acc.maxdd;
      }
    }
  }
  if (!cut) return;
  if (MODE == PM_GEN_NEW) {
    // the later slot loses; several earlier ones may cut it, count it once
    const uint32_t loser = X1 > xj ? X1 : xj;
    if (atomicExch(&a.surv[loser], 0u) != 0u) ++acc.kills;
  } else if (GEN) {
    if (g.inv)
      a.alive_out[xj] = 0u;  // accepted-major: the slot is the proposal (xj its index)
    else
      L.hit[p] = 1u;
  } else {
    if (d2_64 < 0.0) {
      bool w2;
      d2_64 = pair_d2_f64(Pa, Da, Pb, Db, ux, uy, uz, R64, &w2);
    }
    // append to the unsorted list; the host sizes the buffer again and
    // repeats the pass when the count exceeds its capacity
    const uint32_t I = X1 < xj ? X1 : xj;
    const uint32_t J = X1 < xj ? xj : X1;
    const int sgn = X1 < xj ? 1 : -1;
    const uint32_t pos = atomicAdd(a.app_n, 1u);
    if (pos >= a.app_cap) return;
    const double gv = sqrt_cr(d2_64) - ((double)Da.w + (double)Db.w);
    a.outI[pos] = I;
    a.outJ[pos] = J;
    a.outGap[pos] = (float)gv;
    a.outGap64[pos] = gv;
    a.outImg[pos] = pack_img(sgn * ix, sgn * iy, sgn * iz);
  }
}

// Moves entries [base, base + cnt) of the first queue through the screen
// and appends those that need the full test to the second queue; runs the
// full test when 64 are waiting.  Wave-uniform control; the caller holds a
// barrier before the call.
template <int MODE, int PREC>
__device__ __forceinline__ void screen_batch(const PairArgs& a, const CellGeom& g, PairLds& L,
                                             uint32_t base, uint32_t cnt, uint32_t& q2n,
                                             Acc& acc) {
  const int lane = threadIdx.x;
  const bool valid = (uint32_t)lane < cnt;
  const uint32_t e = base + min((uint32_t)lane, cnt - 1u);
  const uint32_t code = screen_test<MODE, PREC>(a, g, L, e, valid, acc);
  const bool f = code != 0u;
  const unsigned long long mask = __ballot(f);
  if (f) L.q2[q2n + lane_prefix(mask)] = L.qE[e] | (code == 2u ? 0x10000u : 0u);
  q2n += (uint32_t)__popcll(mask);
  if (q2n >= (uint32_t)WAVE) {
    __syncthreads();
    full_test<MODE, PREC>(a, g, L, q2n - WAVE + (uint32_t)lane, true, acc);
    q2n -= WAVE;
  }
  __syncthreads();
}

// Walk state over the sorted stencil of one probe chunk.  The stencil is
// taken 64 offsets at a time: every lane looks up one neighbour cell (count,
// start and longest rod in parallel, one round trip) and keeps it in
// registers (cS, cN, cM); the relevant ones form a wave-uniform lane mask,
// taken in stencil (lane) order with readlane.  All other fields are
// wave-uniform.
struct StencilCursor {
  uint32_t sb;              // next stencil block (first offset)
  uint32_t step;            // offsets from one block of this work item to its next
  unsigned long long mask;  // relevant cells of the block not yet taken
  uint32_t cS, cN, cM;      // per lane: the lane's cell of the block
  uint32_t start;           // current neighbour cell: first index, rod count, staged so far
  uint32_t n;
  uint32_t off;
  uint32_t meta;            // image code | same cell << 31
  bool last;                // no stencil block left after the current one
  bool done;
};

template <int MODE>
__device__ void refill_cells(const PairArgs& a, const CellGeom& g, PairLds& L, StencilCursor& c,
                             int c0, int c1, int c2, float chunk_hl, float rr_chunk2) {
  const int lane = threadIdx.x;
  while (!c.last) {
    const uint32_t si = c.sb + (uint32_t)lane;
    const bool valid = si < g.st_n;
    const int4 so = g.st[min(si, g.st_n - 1u)];
    const float md2 = __int_as_float(so.w);
    const int raw[3] = {c0 + so.x, c1 + so.y, c2 + so.z};
    int img[3], m[3];
    // image = floor(raw / nc): the float quotient of raw + 1/2 is at least
    // 1/(2 nc) from an integer and its rounding error is far below that for
    // |raw| < 2^21, so the floor is exact (checked against integer division
    // for every nc <= 1024)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int q = (int)floorf(((float)raw[d] + 0.5f) * g.rnc[d]);
      img[d] = q;
      m[d] = raw[d] - q * g.nc[d];
    }
    const uint32_t nc = g.cell_base + (uint32_t)(m[0] + g.nc[0] * (m[1] + g.nc[1] * m[2]));
    uint32_t n = g.nCount[nc];
    const float hl = g.nHl[nc];
    const uint32_t st0 = g.nStart[nc];
    // the three loads in flight together: the compiler would otherwise sink
    // the second behind a branch on the count and the third past the loop,
    // three dependent round trips instead of one
    asm volatile("" ::"v"(n), "v"(hl), "v"(st0));
    if (MODE == PM_GEN_OLD && !g.inv) n = min(n, g.cap);  // accepted cells hold at most cap rods
    const float rr_c = (chunk_hl + hl) * 1.000001f + g.R_max + CELL_MARGIN;
    const bool beyond = valid && md2 > rr_chunk2;
    const bool rel = valid && !beyond && n > 0u && md2 <= rr_c * rr_c;
    const unsigned long long mk = __ballot(rel);
    const bool same = so.x == 0 && so.y == 0 && so.z == 0;
    c.cS = st0;
    c.cN = n;
    c.cM = (uint32_t)pack_img(img[0], img[1], img[2]) | (same ? 0x80000000u : 0u);
    const bool stop = (__ballot(beyond) != 0ull) || (c.sb + c.step >= g.st_n);
    c.sb += c.step;
    c.mask = mk;
    if (stop) c.last = true;
    if (mk != 0ull) return;
  }
}

// Advances to the next relevant neighbour cell (stencil sorted by minimum
// cell distance: the first block with an offset past the chunk's reach is
// the last one).
template <int MODE>
__device__ __forceinline__ void next_cell(const PairArgs& a, const CellGeom& g, PairLds& L,
                                          StencilCursor& c, int c0, int c1, int c2,
                                          float chunk_hl, float rr_chunk2) {
  if (c.mask == 0ull) {
    if (!c.last) refill_cells<MODE>(a, g, L, c, c0, c1, c2, chunk_hl, rr_chunk2);
    if (c.mask == 0ull) {
      c.done = true;
      return;
    }
  }
  const int i = (int)__builtin_ctzll(c.mask);
  c.mask &= c.mask - 1ull;
  c.start = (uint32_t)__builtin_amdgcn_readlane((int)c.cS, i);
  c.n = (uint32_t)__builtin_amdgcn_readlane((int)c.cN, i);
  c.meta = (uint32_t)__builtin_amdgcn_readlane((int)c.cM, i);
  c.off = 0u;
}

// Prefilter of one (probe, slot) pair, the slot's data in registers:
// centre distance against the pair's own reach bound rp + Q2.w (the probe's
// part from its lane, the slot's staged with the slot), and the
// pass-specific pair filters.  Unfilled slots sit at 1e30, so the
// distance test drops them.  Keeps every pair the full test could count or
// cut (see process_cell).  ORD: the batch holds slots the ordering below
// can drop (else it keeps every pair, and is left out).
template <int MODE, bool ORD>
__device__ __forceinline__ bool prefilter(const float4 P1, float rp, uint32_t X1, uint32_t qprobe,
                                          const float4 Q2, uint32_t xj, uint32_t qa, uint32_t qb) {
  constexpr bool CNT = (MODE == PM_CNT);
  const float wx = P1.x - Q2.x, wy = P1.y - Q2.y, wz = P1.z - Q2.z;
  const float w2 = fmaf(wz, wz, fmaf(wy, wy, wx * wx));
  const float rr = rp + Q2.w;  // the pair's own reach bound
  bool f = w2 <= rr * rr;
  // half-stencil passes (contacts, same-round): a slot of the home cell only
  // after the probe's index; another cell (an image may map back to the home
  // cell) never the probe itself
  if ((CNT || MODE == PM_GEN_NEW) && ORD) f = f && (qprobe < qa || qprobe > qb);
  return f;
}

// Appends the prefilter results of every lane for one probe (both slot
// halves) to the first queue, screening 64 at a time when enough wait.
template <int MODE, int PREC, bool TWO>
__device__ __forceinline__ void push_pairs(const PairArgs& a, const CellGeom& g, PairLds& L,
                                           bool f0, bool f1, uint32_t pp, uint32_t& qn,
                                           uint32_t& q2n, Acc& acc) {
  const int lane = threadIdx.x;
  const unsigned long long m0 = __ballot(f0);
  const uint32_t n0 = (uint32_t)__popcll(m0);
  if (f0) L.qE[qn + lane_prefix(m0)] = pp | ((uint32_t)lane << 8);
  qn += n0;
  if (TWO) {
    const unsigned long long m1 = __ballot(f1);
    if (f1) L.qE[qn + lane_prefix(m1)] = pp | ((uint32_t)(lane + WAVE) << 8);
    qn += (uint32_t)__popcll(m1);
  }
  while (qn >= (uint32_t)WAVE) {
    __syncthreads();
    screen_batch<MODE, PREC>(a, g, L, qn - WAVE, WAVE, q2n, acc);
    qn -= WAVE;
  }
}

// Prefilter of every (probe, slot) pair of one staged batch, looping over
// the probes of the mask (read from their lanes' registers); TWO: the batch
// fills the second slot of the lanes too.  rp_l is the probe part of the
// lane's own probe's reach bound; the slot part sits in Q.w.
template <int MODE, int PREC, bool TWO, bool ORD>
__device__ __forceinline__ void probe_loop(const PairArgs& a, const CellGeom& g, PairLds& L,
                                           unsigned long long pmask, const float4 P1, uint32_t X1,
                                           float rp_l, uint32_t qbase, const float4 Q0, uint32_t x0,
                                           uint32_t qa0, uint32_t qb0, const float4 Q1, uint32_t x1,
                                           uint32_t qa1, uint32_t qb1, uint32_t& qn, uint32_t& q2n,
                                           Acc& acc) {
  for (unsigned long long pm = pmask; pm != 0ull; pm &= pm - 1ull) {
    const int pl = (int)__builtin_ctzll(pm);
    const float4 Pp = make_float4(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(P1.x), pl)),
                                  __int_as_float(__builtin_amdgcn_readlane(__float_as_int(P1.y), pl)),
                                  __int_as_float(__builtin_amdgcn_readlane(__float_as_int(P1.z), pl)),
                                  0.0f);
    const uint32_t Xp = (uint32_t)__builtin_amdgcn_readlane((int)X1, pl);
    const float rp = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rp_l), pl));
    const uint32_t qprobe = qbase + (uint32_t)pl;
    const bool f0 = prefilter<MODE, ORD>(Pp, rp, Xp, qprobe, Q0, x0, qa0, qb0);
    const bool f1 = TWO && prefilter<MODE, ORD>(Pp, rp, Xp, qprobe, Q1, x1, qa1, qb1);
    push_pairs<MODE, PREC, TWO>(a, g, L, f0, f1, (uint32_t)pl, qn, q2n, acc);
  }
}

// One home cell.  For each chunk of up to 64 probes: walk the sorted
// stencil and pack the rods of consecutive relevant neighbour cells into a
// 128-slot staging buffer (two slots per lane, both loads in flight
// together, no lanes wasted on sparse cells); then loop over the chunk's
// probes (read by broadcast from LDS) with the lanes spanning the slots,
// keep the pairs that pass a conservative centre-distance prefilter in an
// LDS queue (ballot + lane prefix), screen them 64 at a time (reach band,
// line distance) and run the full test on the survivors 64 at a time, so
// every expensive step runs on dense work.
//
// Prefilter: the full test first rejects pairs whose FP32 centre distance
// w2 exceeds rr^2 by more than its error band, rr = (hl1 + hl2) * 1.000001
// + R.  The prefilter keeps every pair with w2' <= rr''^2, where
// rr'' = (hl1 * 1.000001 + R_max + 16 ew) * 1.000006 + (hl2 * 1.000001) * 1.000006
// is rr' = hl1 * 1.000001 + hl2 * 1.000001 + R_max + 16 ew >= rr + 16 ew
// scaled by 1.000006 less a few float roundings (below 1e-7 relative), so
// rr''^2 >= rr'^2 * 1.00001; and w' = P1 - (P2 + shift) differs from the
// full test's ((P1 - P2) - shift) by a few roundings of coordinates bounded
// by box + reach, i.e. by less than 6 ew.  So every pair the full test would
// count or cut is queued.  The probe's part of rr'' is computed per lane,
// the slot's part is staged with the slot (Q.w).
//
// A cell is 1 << part_shift work items (parts).  A cell with more than
// split_min probes uses np = min(parts, ceil(probes / split_min)) of them:
// part k < np walks the stencil blocks k, k + np, k + 2 np, ... (64 offsets
// each), so the parts partition the pairs of the cell and a dense cell is
// worked on by several waves at once; parts k >= np return at once.
// The probe bytes, counted per probe, belong to part 0.
template <int MODE, int PREC>
__device__ void process_cell(const PairArgs& a, const CellGeom& g, uint32_t gc, uint32_t part,
                             PairLds& L, Acc& acc) {
  constexpr bool GEN = (MODE == PM_GEN_OLD || MODE == PM_GEN_NEW);
  const int lane = threadIdx.x;
  // an instance without proposals (finished, or planned none) or without
  // accepted rods (its first round) has no pairs: its cells would walk their
  // stencils for nothing
  if (MODE == PM_GEN_OLD && g.idle) return;
  uint32_t pc = uni(g.hCount[gc]);
  if (MODE == PM_GEN_OLD && g.inv) pc = min(pc, g.cap);  // accepted cells hold at most cap rods
  if (pc == 0) return;
  const uint32_t np = min(1u << a.part_shift, (pc + a.split_min - 1u) / a.split_min);
  if (part >= np) return;
  const uint32_t ps = uni(g.hStart[gc]);
  const uint32_t lc = gc - g.cell_base;
  const int c0 = (int)(lc % (uint32_t)g.nc[0]);
  const int c1 = (int)((lc / (uint32_t)g.nc[0]) % (uint32_t)g.nc[1]);
  const int c2 = (int)(lc / ((uint32_t)g.nc[0] * (uint32_t)g.nc[1]));
  const float pad_r = g.R_max + 16.0f * g.ew;

  for (uint32_t pb = 0; pb < pc; pb += WAVE) {
    const uint32_t G = min((uint32_t)WAVE, pc - pb);
    // lane l holds probe l of the chunk (clamped loads, every lane loads)
    const bool inrange = (uint32_t)lane < G;
    const uint32_t pidx = ps + pb + min((uint32_t)lane, G - 1u);
    const float4 P1 = g.hP[pidx];
    const float4 D1 = g.hD[pidx];
    const uint32_t X1 = g.hX[pidx];
    const bool have = inrange;
    L.sP[lane] = P1;
    L.sD[lane] = D1;
    L.sX[lane] = X1;
    L.hit[lane] = 0u;
    // probes taking part, as a wave-uniform mask; the prefilter reads each
    // probe from its lane's registers
    unsigned long long pmask = __ballot(have);
    if (pmask == 0ull) continue;  // every probe of the chunk already cut: nothing to stage
    const float chunk_hl = unif(wave_max(have ? P1.w : 0.0f));
    if (part == 0u) acc.bytes += (unsigned long long)G * 36ull;
    const float rr_chunk = (chunk_hl + g.hl_max) * 1.000001f + g.R_max + CELL_MARGIN;
    const float rr_chunk2 = rr_chunk * rr_chunk;
    uint32_t qn = 0;   // wave-uniform queue lengths: prefiltered pairs,
    uint32_t q2n = 0;  // and pairs waiting for the full test
    StencilCursor cur{part * (uint32_t)WAVE, np * (uint32_t)WAVE, 0ull, 0u, 0u, 0u, 0u, 0u, 0u, 0u,
                      false, false};
    __syncthreads();

    while (!cur.done) {
      // ---- stage: the lane's two slots take their source index from the
      // cell that covers them (registers only)
      uint32_t fill = 0;
      uint32_t src0 = 0u, src1 = 0u, m0 = 0u, m1 = 0u;
      while (fill < (uint32_t)SLOTS) {
        if (cur.off >= cur.n) {
          next_cell<MODE>(a, g, L, cur, c0, c1, c2, chunk_hl, rr_chunk2);
          if (cur.done) break;
        }
        const uint32_t take = min(cur.n - cur.off, (uint32_t)SLOTS - fill);
        const uint32_t r0 = (uint32_t)lane - fill;  // wraps past 2^32 when lane < fill
        const uint32_t r1 = (uint32_t)lane + (uint32_t)WAVE - fill;
        if (r0 < take) {
          src0 = cur.start + cur.off + r0;
          m0 = cur.meta;
        }
        if (r1 < take) {
          src1 = cur.start + cur.off + r1;
          m1 = cur.meta;
        }
        fill += take;
        cur.off += take;
      }
      if (fill == 0) break;
      acc.bytes += (unsigned long long)fill * 36ull;
      // ---- load: both slots' loads in flight together; every lane loads
      // (unfilled slots read index 0 of a non-empty array and are never used)
      const bool v0 = (uint32_t)lane < fill, v1 = (uint32_t)lane + WAVE < fill;
      const float4 p0 = g.nP[src0], p1 = g.nP[src1];
      const float4 d0 = g.nD[src0], d1 = g.nD[src1];
      // accepted-major pass: the slot keeps the proposal's index, which a
      // cut clears in the alive flags
      uint32_t x0 = src0, x1 = src1;
      if (!(MODE == PM_GEN_OLD && g.inv)) {
        x0 = g.nX[src0];
        x1 = g.nX[src1];
      }
      L.bP[lane] = p0;
      L.bD[lane] = d0;
      L.bX[lane] = x0;
      L.bM[lane] = m0;
      L.bP[lane + WAVE] = p1;
      L.bD[lane + WAVE] = d1;
      L.bX[lane + WAVE] = x1;
      L.bM[lane + WAVE] = m1;
      const float3 s0 = shift_f32(g, m0), s1 = shift_f32(g, m1);
      const float far = 1e30f;  // unfilled slots: out of every reach
      // w: the slot's part of the pair reach bound (see the prefilter note)
      const float4 Q0 = v0 ? make_float4(p0.x + s0.x, p0.y + s0.y, p0.z + s0.z, (p0.w * 1.000001f) * 1.000006f)
                           : make_float4(far, far, far, 0.0f);
      const float4 Q1 = v1 ? make_float4(p1.x + s1.x, p1.y + s1.y, p1.z + s1.z, (p1.w * 1.000001f) * 1.000006f)
                           : make_float4(far, far, far, 0.0f);
      // contacts: keep slot iff qprobe < qa || qprobe > qb (home cell: only
      // later indices; elsewhere: anything but the probe itself)
      const uint32_t qa0 = src0, qa1 = src1;
      const uint32_t qb0 = (m0 & 0x80000000u) ? 0xFFFFFFFFu : src0;
      const uint32_t qb1 = (m1 & 0x80000000u) ? 0xFFFFFFFFu : src1;
      // the ordering can only drop a pair whose slot is of the home cell, or
      // is one of the chunk's probes (through an image)
      bool ord = false;
      if (MODE == PM_CNT || MODE == PM_GEN_NEW) {
        const uint32_t qlo = ps + pb;
        const bool o0 = v0 && ((m0 & 0x80000000u) != 0u || src0 - qlo < G);
        const bool o1 = v1 && ((m1 & 0x80000000u) != 0u || src1 - qlo < G);
        ord = __ballot(o0 || o1) != 0ull;
      }
      __syncthreads();
      // ---- prefilter every (probe, slot) pair, queue the survivors
      // each lane's own probe part of the pair reach bound
      const float rp_l = (P1.w * 1.000001f + pad_r) * 1.000006f;
      if (ord) {
        if (fill > (uint32_t)WAVE)
          probe_loop<MODE, PREC, true, true>(a, g, L, pmask, P1, X1, rp_l, ps + pb, Q0, x0, qa0, qb0,
                                             Q1, x1, qa1, qb1, qn, q2n, acc);
        else
          probe_loop<MODE, PREC, false, true>(a, g, L, pmask, P1, X1, rp_l, ps + pb, Q0, x0, qa0, qb0,
                                              Q1, x1, qa1, qb1, qn, q2n, acc);
      } else {
        if (fill > (uint32_t)WAVE)
          probe_loop<MODE, PREC, true, false>(a, g, L, pmask, P1, X1, rp_l, ps + pb, Q0, x0, qa0, qb0,
                                              Q1, x1, qa1, qb1, qn, q2n, acc);
        else
          probe_loop<MODE, PREC, false, false>(a, g, L, pmask, P1, X1, rp_l, ps + pb, Q0, x0, qa0,
                                               qb0, Q1, x1, qa1, qb1, qn, q2n, acc);
      }
      // ---- drain both queues before the slots are restaged
      if (qn > 0) {
        __syncthreads();
        screen_batch<MODE, PREC>(a, g, L, 0u, qn, q2n, acc);
        qn = 0;
      }
      if (q2n > 0) {
        __syncthreads();
        full_test<MODE, PREC>(a, g, L, min((uint32_t)lane, q2n - 1u), (uint32_t)lane < q2n, acc);
        q2n = 0;
      }
      __syncthreads();
      if (MODE == PM_GEN_OLD && !g.inv && !a.test_all) {
        // a proposal once cut is rejected whatever else it overlaps: cut
        // proposals leave the walk, which ends when none is left.  The
        // stencil is sorted by distance, so near a jammed state most
        // proposals leave after their first batches.  The decisions are
        // unchanged; the pair-test count is then the pairs evaluated.
        pmask = __ballot(have && L.hit[lane] == 0u);
        if (pmask == 0ull) break;
      }
    }
    if (MODE == PM_GEN_OLD && !g.inv) {
      // branch-free result store: lanes without a result write a private
      // LDS slot instead
      const bool hitp = have && L.hit[lane] != 0u;
      uint32_t* dst = hitp ? &a.alive_out[pidx] : &L.dummy[lane];
      *dst = 0u;
    }
    __syncthreads();
  }
}

template <int MODE, int PREC>
__global__ void __launch_bounds__(PAIR_THREADS, 4) k_pair(PairArgs a) {
  __shared__ PairLds L;
  constexpr bool GEN = (MODE == PM_GEN_OLD || MODE == PM_GEN_NEW);
  const uint32_t W = a.ncells_total;
  const int xcd = blockIdx.x & 7;
  // rounds after every instance finished plan no candidates: nothing to do
  if (GEN && uni(a.ri->T) == 0u) return;
  // no instance has both candidates and accepted rods: no first-pass pairs
  if (MODE == PM_GEN_OLD && uni(a.ri->old) == 0u) return;
  const bool l0 = threadIdx.x == 0;
  CellGeom g;
  Acc acc = {0u, 0u, 0u, 0.0, 0ull};
  uint32_t cur_inst = 0xFFFFFFFFu;
  if (!GEN) {
    for (int d = 0; d < 3; ++d) {
      g.box[d] = a.box[d];
      g.boxf[d] = a.boxf[d];
      g.nc[d] = a.ncell[d];
      g.rnc[d] = 1.0f / (float)a.ncell[d];
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
    g.inv = false;
    g.idle = false;
  }
  // role pointers of the usual orientation: probes from pr*, staged
  // neighbours from nb*
  g.hP = a.prP;
  g.hD = a.prD;
  g.hX = a.prX;
  g.hStart = a.prStart;
  g.hCount = a.prCount;
  g.nP = a.nbP;
  g.nD = a.nbD;
  g.nX = a.nbX;
  g.nStart = a.nbStart;
  g.nCount = a.nbCount;
  g.nHl = a.nbHl;
  // Lane l watches queue order l: the eight queues of this wave's XCD first
  // (starting at its own sub-queue), then those of the other XCDs.  One
  // vector load of the 64 counters finds the first queue with work left.
  const int sub0 = (blockIdx.x >> 3) & 7;
  const int my_q = (((xcd + (threadIdx.x >> 3)) & 7) << 3) | ((sub0 + (int)threadIdx.x) & 7);
  // queue q: the parts of cells [W q / 64, W (q + 1) / 64) (k_order's ranges)
  const uint32_t my_qb = (uint32_t)(((uint64_t)W * (uint32_t)my_q) / QUEUES_PER_KERNEL) << a.part_shift;
  const uint32_t my_qe = (uint32_t)(((uint64_t)W * (uint32_t)(my_q + 1)) / QUEUES_PER_KERNEL) << a.part_shift;
  for (;;) {
    const uint32_t seen =
        __hip_atomic_load(&a.qctr[my_q * QSTRIDE], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long open = __ballot(seen < my_qe - my_qb);
    if (open == 0ull) break;
    const int pick = (int)__builtin_ctzll(open);  // uniform
    const int q = __builtin_amdgcn_readlane(my_q, pick);
    const uint32_t qb = uni((uint32_t)__builtin_amdgcn_readlane((int)my_qb, pick));
    const uint32_t qe = uni((uint32_t)__builtin_amdgcn_readlane((int)my_qe, pick));
    for (;;) {
      // lane 0 takes the next item; readfirstlane makes it a scalar, so the
      // loop exit below is uniform
      uint32_t old = 0u;
      if (l0) old = atomicAdd(&a.qctr[q * QSTRIDE], 1u);
      const uint32_t item = uni(old);
      if (item >= qe - qb) break;
      const uint32_t ci = (qb + item) >> a.part_shift;
      const uint32_t gc = a.order ? uni(a.order[ci]) : ci;
      const uint32_t part = (qb + item) & ((1u << a.part_shift) - 1u);
      if (GEN) {
        const uint32_t inst = uni(a.cell_inst[gc]);
        if (inst != cur_inst) {
          if (cur_inst != 0xFFFFFFFFu) flush<GEN>(a, g, acc, MODE);
          const InstDesc& d = a.desc[inst];
          for (int dd = 0; dd < 3; ++dd) {
            g.box[dd] = d.smp.box[dd];
            g.boxf[dd] = d.boxf[dd];
            g.nc[dd] = d.ncell[dd];
            g.rnc[dd] = 1.0f / (float)d.ncell[dd];
          }
          g.cell_base = d.cell_base;
          g.st = a.stencil + (MODE == PM_GEN_NEW ? d.hst_off : d.st_off);
          g.st_n = MODE == PM_GEN_NEW ? d.hst_n : d.st_n;
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
          if (MODE == PM_GEN_OLD) {
            // accepted-major when the instance's accepted rods outnumber
            // this round's proposals four to one: fewer, denser probe chunks
            // stage the sparser set (the pair set and every decision are the
            // same).  Below that ratio the proposals are the probes, so that
            // cut ones can leave their walk early.
            const uint32_t nacc = uni(a.st[inst].n_acc);
            const uint32_t nprop = uni(a.prStart[d.cell_base + d.ncells]) - uni(a.prStart[d.cell_base]);
            g.inv = nacc > 4u * nprop;
            g.idle = nprop == 0u || nacc == 0u;
            g.hP = g.inv ? a.nbP : a.prP;
            g.hD = g.inv ? a.nbD : a.prD;
            g.hX = g.inv ? a.nbX : a.prX;
            g.hStart = g.inv ? a.nbStart : a.prStart;
            g.hCount = g.inv ? a.nbCount : a.prCount;
            g.nP = g.inv ? a.prP : a.nbP;
            g.nD = g.inv ? a.prD : a.nbD;
            g.nX = g.inv ? a.prX : a.nbX;
            g.nStart = g.inv ? a.prStart : a.nbStart;
            g.nCount = g.inv ? a.prCount : a.nbCount;
            g.nHl = g.inv ? a.prHl : a.nbHl;
          } else {
            g.inv = false;
            g.idle = false;
          }
        }
      }
      process_cell<MODE, PREC>(a, g, gc, part, L, acc);
    }
  }
  if (!GEN || cur_inst != 0xFFFFFFFFu) flush<GEN>(a, g, acc, MODE);
}

// Work order of a pair pass: within each of the 64 queue ranges of cells,
// the cells by decreasing estimated cost, so that each queue starts with its
// costliest cells and the launch ends on light ones.  The cost of a cell is
// the rods around (its probe and neighbour counts) times the volume its
// stencil walk covers, (h + H)^3, h the cell's longest rod (half length) and
// H the instance's longest half length plus a cell edge; two buckets per
// octave.  One workgroup per range; the order inside a bucket is arbitrary,
// as every result of a pair pass is independent of the order of its work.
constexpr int ORDER_BUCKETS = 32;
constexpr int ORDER_THREADS = 256;
__device__ __forceinline__ uint32_t order_bucket(const PairArgs& a, uint32_t c) {
  const InstDesc& d = a.desc[a.cell_inst[c]];
  const float h = fmaxf(a.prHl[c], a.nbHl[c]) + d.hl_max + d.edge[0];
  const float w = (float)(a.prCount[c] + a.nbCount[c]) * h * h * h;
  // (bits >> 22) = 2 * biased exponent + first mantissa bit; bucket 0: w >= 2^13
  const int k = 283 - (int)(__float_as_uint(w) >> 22);
  return (uint32_t)min(max(k, 0), ORDER_BUCKETS - 1);
}

__global__ void __launch_bounds__(ORDER_THREADS) k_order(PairArgs a) {
  __shared__ uint32_t slot[ORDER_BUCKETS];
  const uint32_t r = blockIdx.x;
  const uint32_t c0 = (uint32_t)(((uint64_t)a.ncells_total * r) / QUEUES_PER_KERNEL);
  const uint32_t c1 = (uint32_t)(((uint64_t)a.ncells_total * (r + 1u)) / QUEUES_PER_KERNEL);
  if (threadIdx.x < ORDER_BUCKETS) slot[threadIdx.x] = 0u;
  __syncthreads();
  for (uint32_t c = c0 + threadIdx.x; c < c1; c += ORDER_THREADS) atomicAdd(&slot[order_bucket(a, c)], 1u);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t sum = c0;
    for (int k = 0; k < ORDER_BUCKETS; ++k) {
      const uint32_t t = slot[k];
      slot[k] = sum;
      sum += t;
    }
  }
  __syncthreads();
  for (uint32_t c = c0 + threadIdx.x; c < c1; c += ORDER_THREADS)
    a.order[atomicAdd(&slot[order_bucket(a, c)], 1u)] = c;
}

static int g_cus = 0;

// Two parts per cell for the reject-vs-accepted pass, whose
// agglomerate-dense cells otherwise end a launch alone; none for the
// same-round and contact passes.  Used only when a launch has fewer than 16
// cells per wave (launch_pair): with more, dense cells are spread over many
// items per wave anyway, and the grabs of empty parts cost more than they
// save.  Measured on MI355X (500k-rod generation, 12k cells, with the
// dense-first order of k_order): reject-vs-accepted pass 2.72 ms with two
// parts of at least 32 probes, 2.77 with 24, 2.97 with four parts;
// same-round pass 1.80 ms unsplit, 1.91-2.06 split.  The 256-packing batch
// (1M+ cells): 301 ms unsplit, 312 split.
int pair_part_shift(int mode) { return mode == PM_GEN_OLD ? 1 : 0; }

int pair_split_min(int mode) {
  (void)mode;
  return 32;
}

int pair_grid(int mode) {
  if (g_cus == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    g_cus = cus;
  }
  (void)mode;
  return g_cus * 16;
}

void launch_pair(int mode, int prec, const PairArgs& args, hipStream_t s, const PairTiles* t) {
  // the handle's tuned launch shape where set (nanofiller_amd/autotune.py)
  uint32_t waves = 0;
  if (t) waves = mode == PM_GEN_OLD ? t->waves_old : (mode == PM_GEN_NEW ? t->waves_new : 0u);
  const int cus = pair_grid(mode) / 16;
  const dim3 grid(waves ? (uint32_t)cus * waves : (uint32_t)pair_grid(mode)), block(PAIR_THREADS);
  PairArgs a = args;
  // launches with few cells per wave: costliest cells first (under 64 per
  // wave), dense cells split (under 16 per wave)
  const bool small = a.ncells_total < 16u * grid.x;
  const int ps = (t && mode == PM_GEN_OLD && t->part_shift >= 0) ? t->part_shift : pair_part_shift(mode);
  a.part_shift = small ? (uint32_t)ps : 0u;
  if (a.ncells_total < 64u * grid.x && a.order != nullptr)
    hipLaunchKernelGGL(k_order, dim3(QUEUES_PER_KERNEL), dim3(ORDER_THREADS), 0, s, a);
  else
    a.order = nullptr;
  a.split_min = (t && t->split_min) ? t->split_min : (uint32_t)pair_split_min(mode);
  a.test_all = (t && mode == PM_GEN_OLD) ? t->test_all : 0u;
#define NF_LAUNCH(M, P) hipLaunchKernelGGL((k_pair<M, P>), grid, block, 0, s, a)
  if (prec == NF_PREC_FP64) {
    switch (mode) {
      case PM_GEN_OLD: NF_LAUNCH(PM_GEN_OLD, NF_PREC_FP64); break;
      case PM_GEN_NEW: NF_LAUNCH(PM_GEN_NEW, NF_PREC_FP64); break;
      default: NF_LAUNCH(PM_CNT, NF_PREC_FP64); break;
    }
  } else {
    switch (mode) {
      case PM_GEN_OLD: NF_LAUNCH(PM_GEN_OLD, NF_PREC_FP32_RECHECK); break;
      case PM_GEN_NEW: NF_LAUNCH(PM_GEN_NEW, NF_PREC_FP32_RECHECK); break;
      default: NF_LAUNCH(PM_CNT, NF_PREC_FP32_RECHECK); break;
    }
  }
#undef NF_LAUNCH
}

}  // namespace nf
