// nf_gen.hip - one insert/reject round for a batch of packings:
//   ctl_begin -> gen -> bin (scan + scatter) -> reject_old -> reject_new
//   -> rank (scan) -> commit -> ctl_end.
// Every launch has a fixed shape and reads its work size from device
// memory, so a round is captured once in a HIP graph and replayed with no
// host round trip.  Replaces the host-driven round of cudaGenStep
// (src/cudaMain.cu:219-351) and its per-round device-to-host sync
// (checkArray, src/cudaMain.cu:201-216).
#include "nf_internal.h"

namespace nf {

constexpr int GEN_THREADS = 256;

// ---------------------------------------------------------------------------
// Round schedule: candidate slots per round as an integer function of the
// round history (same formula as the oracle).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t clamp_batch(uint64_t m, const InstDesc& d) {
  if (m < d.min_batch) m = d.min_batch;
  if (m > d.max_batch) m = d.max_batch;
  return (uint32_t)m;
}

__device__ __forceinline__ uint32_t schedule(const InstDesc& d, const InstState& s) {
  const uint32_t deficit = d.n_target - s.n_acc;
  if (s.round == 0) return clamp_batch((uint64_t)deficit + (uint64_t)(deficit / 4u), d);
  if (s.aprev == 0) return d.max_batch;
  const uint64_t num = (uint64_t)deficit * (uint64_t)s.mprev * 5ull;
  const uint64_t den = (uint64_t)s.aprev * 4ull;
  return clamp_batch((num + den - 1) / den, d);
}

// Exclusive scan over the 64 lanes of one wave (all lanes active).
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* total) {
  const int lane = threadIdx.x & (WAVE - 1);
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, WAVE);
    inc += (lane >= o) ? t : 0u;
  }
  *total = (uint32_t)__builtin_amdgcn_readlane((int)inc, WAVE - 1);
  return inc - v;
}

// Plans the round: M per running instance, slot prefix, queue reset (the
// first wave; lane l owns a contiguous chunk of instances), and clears the
// per-cell proposal counts and longest proposals (the whole grid).
__global__ void __launch_bounds__(GEN_THREADS) k_ctl_begin(const InstDesc* __restrict__ desc,
                                                           InstState* __restrict__ st,
                                                           RoundInfo* __restrict__ ri,
                                                           uint32_t* __restrict__ slot_base,
                                                           uint32_t* __restrict__ queues, int n_inst,
                                                           uint32_t* __restrict__ pcount,
                                                           uint32_t* __restrict__ phl,
                                                           uint64_t ncells) {
  // every thread: clear the per-cell proposal counts and longest proposals
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncells;
       c += (uint64_t)gridDim.x * blockDim.x) {
    pcount[c] = 0u;
    phl[c] = 0u;
  }
  // the first wave: the controller
  if (blockIdx.x != 0 || threadIdx.x >= WAVE) return;
  const int lane = threadIdx.x;
  const int chunk = (n_inst + WAVE - 1) / WAVE;
  const int b = lane * chunk;
  const int e = min(b + chunk, n_inst);
  uint32_t sum = 0, old = 0;
  for (int i = b; i < e; ++i) {
    uint32_t M = 0;
    if (st[i].status == NF_STATUS_RUNNING) M = schedule(desc[i], st[i]);
    st[i].M = M;
    slot_base[i] = sum;
    sum += M;
    old += (M > 0u && st[i].n_acc > 0u) ? 1u : 0u;
  }
  uint32_t total, old_total;
  const uint32_t off = wave_excl_scan(sum, &total);
  (void)wave_excl_scan(old, &old_total);
  for (int i = b; i < e; ++i) slot_base[i] += off;
  if (lane == 0) {
    slot_base[n_inst] = total;
    ri->T = total;
    ri->old = old_total;
  }
  for (int q = lane; q < Q_N * QUEUES_PER_KERNEL; q += WAVE) queues[q * QSTRIDE] = 0;
}

__device__ __forceinline__ int find_inst(const uint32_t* __restrict__ base, int n, uint32_t g) {
  int lo = 0, hi = n;  // base[lo] <= g < base[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (base[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t cell_of(const InstDesc& d, const float4 P) {
  int c[3];
  const float p[3] = {P.x, P.y, P.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int i = (int)((double)p[a] * d.cell_scale[a]);
    i = i < 0 ? 0 : (i >= d.ncell[a] ? d.ncell[a] - 1 : i);
    c[a] = i;
  }
  return (uint32_t)(c[0] + d.ncell[0] * (c[1] + d.ncell[1] * c[2]));
}

// Draws the candidates of every slot [0, T); bins proposals by cell.
__global__ void __launch_bounds__(GEN_THREADS) k_gen(const InstDesc* __restrict__ desc,
                                                     const InstState* __restrict__ st,
                                                     const RoundInfo* __restrict__ ri,
                                                     const uint32_t* __restrict__ slot_base,
                                                     int n_inst, uint64_t T_cap,
                                                     float4* __restrict__ sP, float4* __restrict__ sD,
                                                     int32_t* __restrict__ scell,
                                                     uint32_t* __restrict__ srk,
                                                     uint32_t* __restrict__ surv,
                                                     uint32_t* __restrict__ palive,
                                                     uint32_t* __restrict__ pcount,
                                                     float* __restrict__ phl) {
  const uint32_t T = ri->T;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g <= T_cap;
       g += (uint64_t)gridDim.x * blockDim.x) {
    surv[g] = 0;
    if (g < T_cap) palive[g] = 0;  // the survivors' scan reads all T_cap flags
    if (g >= T) continue;
    const int inst = find_inst(slot_base, n_inst, (uint32_t)g);
    const InstDesc& d = desc[inst];
    const uint32_t slot = (uint32_t)g - slot_base[inst];
    Rod r;
    const int v = sample_candidate(d.smp, st[inst].round, slot, &r);
    if (v == 1) {
      const uint32_t gc = d.cell_base + cell_of(d, r.P);
      srk[g] = atomicAdd(&pcount[gc], 1u);
      atomicMax((unsigned int*)&phl[gc], __float_as_uint(r.P.w));
      sP[g] = r.P;
      sD[g] = r.D;
      scell[g] = (int32_t)gc;
    } else {
      scell[g] = -1;
    }
  }
}

// Moves proposals into their cell ranges.
__global__ void __launch_bounds__(GEN_THREADS) k_scatter(const RoundInfo* __restrict__ ri,
                                                         const float4* __restrict__ sP,
                                                         const float4* __restrict__ sD,
                                                         const int32_t* __restrict__ scell,
                                                         const uint32_t* __restrict__ srk,
                                                         const uint32_t* __restrict__ pstart,
                                                         float4* __restrict__ pP,
                                                         float4* __restrict__ pD,
                                                         uint32_t* __restrict__ pslot,
                                                         uint32_t* __restrict__ palive,
                                                         uint32_t* __restrict__ surv) {
  const uint32_t T = ri->T;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < T; g += gridDim.x * blockDim.x) {
    const int32_t c = scell[g];
    if (c < 0) continue;
    const uint32_t dst = pstart[c] + srk[g];
    pP[dst] = sP[g];
    pD[dst] = sD[g];
    pslot[dst] = g;
    palive[dst] = 1u;
    surv[g] = 1u;  // cleared for the proposals that either pass rejects
  }
}

// Moves the survivors of the reject-vs-accepted pass to their compacted
// positions - aidx is the exclusive prefix of the alive flags over the
// cell-sorted proposals, so each cell's survivors stay together and in
// order - sets the compacted cell ranges, and clears the survivor flag of
// the proposals that pass cut.
__global__ void __launch_bounds__(GEN_THREADS) k_alive_scatter(const uint32_t* __restrict__ pstart,
                                                               const float4* __restrict__ pP,
                                                               const float4* __restrict__ pD,
                                                               const uint32_t* __restrict__ pslot,
                                                               const uint32_t* __restrict__ palive,
                                                               const uint32_t* __restrict__ aidx,
                                                               float4* __restrict__ aP,
                                                               float4* __restrict__ aD,
                                                               uint32_t* __restrict__ aslot,
                                                               uint32_t* __restrict__ astart,
                                                               uint32_t* __restrict__ acount,
                                                               uint32_t* __restrict__ surv,
                                                               uint64_t ncells) {
  const uint64_t P = pstart[ncells];  // proposals of this round, all instances
  const uint64_t n = P > ncells + 1 ? P : ncells + 1;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (uint64_t)gridDim.x * blockDim.x) {
    if (k < P) {
      const uint32_t g = pslot[k];
      if (palive[k]) {
        const uint32_t dst = aidx[k];
        aP[dst] = pP[k];
        aD[dst] = pD[k];
        aslot[dst] = g;
      } else {
        surv[g] = 0u;
      }
    }
    if (k < ncells) {
      const uint32_t a0 = aidx[pstart[k]];
      astart[k] = a0;
      acount[k] = aidx[pstart[k + 1]] - a0;
    } else if (k == ncells) {
      astart[k] = aidx[pstart[k]];
    }
  }
}

// Accepts survivors in slot order up to the target; inserts them by id and
// into their cell slots.
__global__ void __launch_bounds__(GEN_THREADS) k_commit(const InstDesc* __restrict__ desc,
                                                        InstState* __restrict__ st,
                                                        const RoundInfo* __restrict__ ri,
                                                        const uint32_t* __restrict__ slot_base,
                                                        int n_inst,
                                                        const float4* __restrict__ sP,
                                                        const float4* __restrict__ sD,
                                                        const int32_t* __restrict__ scell,
                                                        const uint32_t* __restrict__ surv,
                                                        const uint32_t* __restrict__ srank,
                                                        float4* __restrict__ bP,
                                                        float4* __restrict__ bD,
                                                        const uint32_t* __restrict__ ostart,
                                                        uint32_t* __restrict__ ocount,
                                                        float* __restrict__ ohl,
                                                        float4* __restrict__ oP,
                                                        float4* __restrict__ oD,
                                                        uint32_t* __restrict__ oid) {
  const uint32_t T = ri->T;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < T; g += gridDim.x * blockDim.x) {
    if (!surv[g]) continue;
    const int inst = find_inst(slot_base, n_inst, g);
    const InstDesc& d = desc[inst];
    const uint32_t n_acc = st[inst].n_acc;
    const uint32_t r = srank[g] - srank[slot_base[inst]];
    if (r >= d.n_target - n_acc) continue;
    const uint32_t id = n_acc + r;
    const float4 P = sP[g], D = sD[g];
    bP[d.rod_off + id] = P;
    bD[d.rod_off + id] = D;
    const int32_t c = scell[g];
    const uint32_t k = atomicAdd(&ocount[c], 1u);
    if (k < d.cap) {
      const uint32_t dst = ostart[c] + k;
      oP[dst] = P;
      oD[dst] = D;
      oid[dst] = id;
      atomicMax((unsigned int*)&ohl[c], __float_as_uint(P.w));
    } else {
      // no slot: give the count back, so that once this kernel ends every
      // cell's count is at most its capacity (later passes of the other
      // instances of a batch read a cell's first `count` slots), and stop
      // the instance (the host generates it again with more slots)
      atomicSub(&ocount[c], 1u);
      atomicOr(&st[inst].overflow, 1u);
    }
  }
}

// Closes the round: accepted counts, statistics, termination.  One wave.
__global__ void __launch_bounds__(WAVE) k_ctl_end(const InstDesc* __restrict__ desc,
                                                  InstState* __restrict__ st,
                                                  RoundInfo* __restrict__ ri,
                                                  const uint32_t* __restrict__ slot_base,
                                                  const uint32_t* __restrict__ srank,
                                                  const uint32_t* __restrict__ pstart,
                                                  const uint32_t* __restrict__ astart, int n_inst) {
  const int lane = threadIdx.x;
  uint32_t running = 0, est = 0;
  for (int i = lane; i < n_inst; i += WAVE) {
    InstState s = st[i];
    if (s.status != NF_STATUS_RUNNING) continue;
    const InstDesc& d = desc[i];
    const uint32_t sb = slot_base[i];
    const uint32_t surv = srank[sb + s.M] - srank[sb];
    const uint32_t deficit = d.n_target - s.n_acc;
    const uint32_t acc = surv < deficit ? surv : deficit;
    const uint32_t P = pstart[d.cell_base + d.ncells] - pstart[d.cell_base];
    const uint32_t A = astart[d.cell_base + d.ncells] - astart[d.cell_base];
    s.candidates += s.M;
    s.proposals += P;
    s.rej_old += P - A;  // cut by an accepted rod
    st[i].rej_old = s.rej_old;
    s.n_acc += acc;
    s.round += 1;
    s.mprev = s.M;
    s.aprev = acc;
    s.stall = acc == 0 ? s.stall + 1 : 0;
    if (s.overflow)
      s.status = NF_STATUS_OVERFLOW;
    else if (s.n_acc == d.n_target)
      s.status = NF_STATUS_TARGET;
    else if (s.stall >= d.stall_rounds)
      s.status = NF_STATUS_STALLED;
    else if (s.round >= d.max_rounds)
      s.status = NF_STATUS_MAX_ROUNDS;
    st[i].candidates = s.candidates;
    st[i].proposals = s.proposals;
    st[i].n_acc = s.n_acc;
    st[i].round = s.round;
    st[i].mprev = s.mprev;
    st[i].aprev = s.aprev;
    st[i].stall = s.stall;
    st[i].status = s.status;
    running += (s.status == NF_STATUS_RUNNING) ? 1u : 0u;
    if (s.status == NF_STATUS_RUNNING) {
      const uint32_t left = d.n_target - s.n_acc;
      const uint32_t e = acc ? (left + acc - 1u) / acc : 1u;
      est = e > est ? e : est;
    }
  }
  uint32_t v = running;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v += __shfl_xor(v, o, WAVE);
    const uint32_t w = __shfl_xor(est, o, WAVE);
    est = w > est ? w : est;
  }
  if (lane == 0) {
    ri->running = v;
    ri->rounds += 1;
    ri->est = est;
  }
}

static int gen_grid(uint64_t n) {
  uint64_t b = (n + GEN_THREADS - 1) / GEN_THREADS;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

// One round.  ev (optional, 9 events) brackets the kernel classes.
void launch_round(GenBuffers& b, int prec, int diag, hipStream_t s, hipEvent_t* ev) {
  const int n = b.n_inst;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(k_ctl_begin, dim3(gen_grid(b.total_cells)), dim3(GEN_THREADS), 0, s, b.desc,
                     b.st, b.ri, b.slot_base, b.queues, n, b.pcount, (uint32_t*)b.phl,
                     (uint64_t)b.total_cells);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_gen, dim3(gen_grid(b.T_cap + 1)), dim3(GEN_THREADS), 0, s, b.desc, b.st,
                     b.ri, b.slot_base, n, b.T_cap, b.sP, b.sD, b.scell, b.srk, b.surv, b.palive,
                     b.pcount, b.phl);
  if (ev) (void)hipEventRecord(ev[2], s);
  scan_u32(b.pcount, b.pstart, b.total_cells, b.scan_tmp, s);
  hipLaunchKernelGGL(k_scatter, dim3(gen_grid(b.T_cap)), dim3(GEN_THREADS), 0, s, b.ri, b.sP,
                     b.sD, b.scell, b.srk, b.pstart, b.pP, b.pD, b.pslot, b.palive, b.surv);
  if (ev) (void)hipEventRecord(ev[3], s);

  PairArgs a{};
  a.desc = b.desc;
  a.ri = b.ri;
  a.st = b.st;
  a.cell_inst = b.cell_inst;
  a.ncells_total = (uint32_t)b.total_cells;
  a.stencil = b.stencil;
  a.diag = diag;
  a.order = b.order;
  // stage 1: proposals vs accepted rods
  a.prP = b.pP;
  a.prD = b.pD;
  a.prX = b.pslot;
  a.prStart = b.pstart;
  a.prCount = b.pcount;
  a.prHl = b.phl;
  a.nbP = b.oP;
  a.nbD = b.oD;
  a.nbX = b.oid;
  a.nbStart = b.ostart;
  a.nbCount = b.ocount;
  a.nbHl = b.ohl;
  a.qctr = b.queues + Q_OLD * QWORDS;
  a.alive_out = b.palive;
  launch_pair(PM_GEN_OLD, prec, a, s, &b.tiles);
  if (ev) (void)hipEventRecord(ev[4], s);
  // stage 2: survivors vs higher-priority survivors, over the survivors
  // compacted per cell (a cell's longest rod stays that of all its
  // proposals: a bound, as the stencil walk needs)
  scan_u32(b.palive, b.aidx, b.T_cap, b.scan_tmp, s);
  hipLaunchKernelGGL(k_alive_scatter, dim3(gen_grid(std::max(b.T_cap, b.total_cells + 1))),
                     dim3(GEN_THREADS), 0, s, b.pstart, b.pP, b.pD, b.pslot, b.palive, b.aidx, b.aP,
                     b.aD, b.aslot, b.astart, b.acount, b.surv, (uint64_t)b.total_cells);
  a.prP = b.aP;
  a.prD = b.aD;
  a.prX = b.aslot;
  a.prStart = b.astart;
  a.prCount = b.acount;
  a.prHl = b.phl;
  a.nbP = b.aP;
  a.nbD = b.aD;
  a.nbX = b.aslot;
  a.nbStart = b.astart;
  a.nbCount = b.acount;
  a.nbHl = b.phl;
  a.qctr = b.queues + Q_NEW * QWORDS;
  a.alive_out = nullptr;
  a.surv = b.surv;
  launch_pair(PM_GEN_NEW, prec, a, s, &b.tiles);
  if (ev) (void)hipEventRecord(ev[5], s);
  scan_u32(b.surv, b.srank, b.T_cap, b.scan_tmp, s);
  hipLaunchKernelGGL(k_commit, dim3(gen_grid(b.T_cap)), dim3(GEN_THREADS), 0, s, b.desc, b.st,
                     b.ri, b.slot_base, n, b.sP, b.sD, b.scell, b.surv, b.srank, b.bP, b.bD,
                     b.ostart, b.ocount, b.ohl, b.oP, b.oD, b.oid);
  if (ev) (void)hipEventRecord(ev[6], s);
  hipLaunchKernelGGL(k_ctl_end, dim3(1), dim3(WAVE), 0, s, b.desc, b.st, b.ri,
                     b.slot_base, b.srank, b.pstart, b.astart, n);
  if (ev) (void)hipEventRecord(ev[7], s);
}

}  // namespace nf
