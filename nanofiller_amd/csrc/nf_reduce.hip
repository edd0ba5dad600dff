// nf_reduce.hip - dead-end pruning of the contact network: the `reduce` verb
// of the reference (src/NanoFiller.cpp:71-90) driving reduce()
// (src/conductivity.cu:79-138) with contact_count_kernel_full and
// reductKernel (inc/conductKernels.h:40-162).
//
// Rule: rounds of simultaneous removal.  In each round every rod still
// present whose contact count among the rods still present is below
// min_cont is removed; the rounds stop when one removes nothing.  Contacts
// are those of the last nf_contacts call (each periodic image of a pair is
// one contact, as the reference's stencil sweep counts it) except those
// whose image crosses the electrode axis, which is not periodic for the
// network.  The reference keeps every rod of its electrode cell layers
// (is_internal in reductKernel); here the electrodes are the planes of the
// network (nf_network), and a rod's coupling to a plane counts as one
// contact, so a rod that bridges an electrode and one other rod stays.
//
// Device layout: state[i] (0 present, 1 removed this round, 2 removed
// earlier), deg[i] (contacts among present rods plus couplings), one
// removal counter per round.  A round is two kernels: mark (retire the last
// round's removals, remove the rods now below min_cont) and propagate
// (every contact between a rod removed this round and a present rod takes
// one from the present rod's count).  The host queues rounds in batches and
// reads the counters between batches; a round that removes nothing sets a
// stop flag that makes the rest of the batch exit at once.  The kept rods
// are compacted in id order (scan of the kept flags, scatter) and replace
// the packing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "nf_host.h"

using namespace nf;

namespace {

constexpr int RED_THREADS = 256;
constexpr int RED_BATCH = 16;  // rounds queued between host reads

struct RedArgs {
  uint32_t n;
  const uint32_t* rowptr;
  const uint32_t* J;
  const int32_t* img;
  const float4* P;
  const float4* D;
  int axis;
  double L, gcut;
  uint32_t min_cont;
  uint32_t* state;
  uint32_t* deg;
  uint32_t* cnt;   // RED_BATCH removal counters
  uint32_t* stop;  // 1 once a round removed nothing
};

__device__ __forceinline__ int red_img(int32_t v, int axis) {
  return (int)(int8_t)((v >> (8 * axis)) & 0xFF);
}

uint32_t red_grid(uint32_t n) {
  const uint32_t g = (n + RED_THREADS - 1) / RED_THREADS;
  return g < 1 ? 1 : (g > 4096 ? 4096 : g);
}

// Couplings to the electrode planes at 0 and L along the axis (the rule of
// nf_network): the capsule's surface gap to the plane within the cut.
__global__ void k_red_init(RedArgs a) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
    const float4 p = a.P[i], d = a.D[i];
    const float pa = a.axis == 0 ? p.x : (a.axis == 1 ? p.y : p.z);
    const float ca = a.axis == 0 ? d.x : (a.axis == 1 ? d.y : d.z);
    const double ext = fabs((double)p.w * (double)ca);
    const double pos = (double)pa, rad = (double)d.w;
    const double gs = (pos - ext) - rad;
    const double gk = (a.L - (pos + ext)) - rad;
    a.deg[i] = (gs <= a.gcut ? 1u : 0u) + (gk <= a.gcut ? 1u : 0u);
    a.state[i] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.stop = 0u;
}

// Contact counts: both ends of every contact that does not cross the axis.
__global__ void k_red_degree(RedArgs a) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.n; r += gridDim.x * blockDim.x) {
    const uint32_t e = a.rowptr[r + 1];
    uint32_t own = 0;
    for (uint32_t k = a.rowptr[r]; k < e; ++k) {
      if (red_img(a.img[k], a.axis) != 0) continue;
      ++own;
      atomicAdd(&a.deg[a.J[k]], 1u);
    }
    if (own) atomicAdd(&a.deg[r], own);
  }
}

__global__ void k_red_mark(RedArgs a, uint32_t slot) {
  if (__hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  uint32_t local = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
    uint32_t s = a.state[i];
    if (s == 1u) {
      s = 2u;
      a.state[i] = 2u;
    }
    if (s == 0u && a.deg[i] < a.min_cont) {
      a.state[i] = 1u;
      ++local;
    }
  }
  // one atomic per wave after the loop (all lanes converged)
  const unsigned long long w = wave_sum_u64(local);
  if ((threadIdx.x & (WAVE - 1)) == 0 && w) atomicAdd(&a.cnt[slot], (uint32_t)w);
}

__global__ void k_red_propagate(RedArgs a, uint32_t slot) {
  if (__hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  if (a.cnt[slot] == 0u) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.stop = 1u;
    return;
  }
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.n; r += gridDim.x * blockDim.x) {
    const uint32_t e = a.rowptr[r + 1];
    const uint32_t sr = a.state[r];
    if (sr == 2u) continue;
    uint32_t lost = 0;
    for (uint32_t k = a.rowptr[r]; k < e; ++k) {
      if (red_img(a.img[k], a.axis) != 0) continue;
      const uint32_t j = a.J[k];
      const uint32_t sj = a.state[j];
      if (sr == 1u && sj == 0u) atomicSub(&a.deg[j], 1u);
      if (sr == 0u && sj == 1u) ++lost;
    }
    if (lost) atomicSub(&a.deg[r], lost);
  }
}

__global__ void k_red_keep(const uint32_t* __restrict__ state, uint32_t n, uint32_t* __restrict__ keep) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    keep[i] = state[i] == 0u ? 1u : 0u;
}

__global__ void k_red_scatter(const uint32_t* __restrict__ state, const uint32_t* __restrict__ pos,
                              uint32_t n, const float4* __restrict__ P, const float4* __restrict__ D,
                              float4* __restrict__ oP, float4* __restrict__ oD,
                              uint32_t* __restrict__ oid) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (state[i] != 0u) continue;
    const uint32_t q = pos[i];
    oP[q] = P[i];
    oD[q] = D[i];
    oid[q] = i;
  }
}

}  // namespace

extern "C" int nf_reduce(nf_handle* h, int32_t min_cont, nf_reduce_info* info) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  h->err.clear();
  if (min_cont < 0) return fail(h, NF_E_INVALID, "min_cont must be >= 0");
  ContactBuffers& c = h->cb;
  if (!c.valid) return fail(h, NF_E_STATE, "nf_reduce needs nf_contacts first");
  const nf_params& p = h->params[h->post_inst];
  const uint32_t n = c.n_rods;
  const uint32_t off = h->desc[h->post_inst].rod_off;
  float4* P = h->gb.bP + off;
  float4* D = h->gb.bD + off;
  hipStream_t s = h->stream;
  RedArgs a{};
  a.n = n;
  a.rowptr = c.rowptr;
  a.J = c.J;
  a.img = c.img;
  a.P = P;
  a.D = D;
  a.axis = p.elec_axis;
  a.L = p.box[p.elec_axis];
  a.gcut = p.contact_gap;  // the electrode rule of nf_network
  a.min_cont = (uint32_t)min_cont;
  uint32_t *keep = nullptr, *pos = nullptr, *tmp = nullptr, *oid = nullptr;
  float4 *oP = nullptr, *oD = nullptr;
  const uint64_t n1 = (uint64_t)n + 1;
  int rc = dalloc(h, &a.state, n1);
  if (rc == NF_OK) rc = dalloc(h, &a.deg, n1);
  if (rc == NF_OK) rc = dalloc(h, &a.cnt, RED_BATCH + 1);
  if (rc == NF_OK) rc = dalloc(h, &keep, n1);
  if (rc == NF_OK) rc = dalloc(h, &pos, n1);
  if (rc == NF_OK) rc = dalloc(h, &tmp, scan_tmp_words(n1));
  if (rc == NF_OK) rc = dalloc(h, &oid, n1);
  if (rc == NF_OK) rc = dalloc(h, &oP, n1);
  if (rc == NF_OK) rc = dalloc(h, &oD, n1);
  a.stop = a.cnt + RED_BATCH;
  auto release = [&]() {
    void* w[] = {a.state, a.deg, a.cnt, keep, pos, tmp, oid, oP, oD};
    for (void* q : w) dfree(h, q);
  };
  if (rc != NF_OK) {
    release();
    return rc;
  }
  hipEvent_t* ev = h->ev;
  hipEventRecord(ev[0], s);
  const dim3 grid(red_grid(n)), block(RED_THREADS);
  hipLaunchKernelGGL(k_red_init, grid, block, 0, s, a);
  hipLaunchKernelGGL(k_red_degree, grid, block, 0, s, a);
  uint32_t rounds = 0;
  bool done = false;
  std::vector<uint32_t> hc(RED_BATCH);
  hipError_t e = hipGetLastError();
  // every round but the last removes at least one rod: at most n + 1 rounds
  while (e == hipSuccess && !done && rounds <= n) {
    fill_u32(a.cnt, 0u, RED_BATCH, s);
    for (uint32_t b = 0; b < (uint32_t)RED_BATCH; ++b) {
      hipLaunchKernelGGL(k_red_mark, grid, block, 0, s, a, b);
      hipLaunchKernelGGL(k_red_propagate, grid, block, 0, s, a, b);
    }
    e = hipMemcpyAsync(hc.data(), a.cnt, sizeof(uint32_t) * RED_BATCH, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = sync_stream(s, "reduce rounds");
    if (e == hipSuccess) e = hipGetLastError();
    for (uint32_t b = 0; b < (uint32_t)RED_BATCH && e == hipSuccess; ++b) {
      ++rounds;
      if (hc[b] == 0u) {
        done = true;
        break;
      }
    }
  }
  if (e == hipSuccess && !done) {
    release();
    return fail(h, NF_E_STATE, "reduce did not settle");
  }
  uint32_t kept = 0;
  if (e == hipSuccess && n) {
    hipLaunchKernelGGL(k_red_keep, grid, block, 0, s, a.state, n, keep);
    scan_u32(keep, pos, n, tmp, s);  // pos[n] = kept
    hipLaunchKernelGGL(k_red_scatter, grid, block, 0, s, a.state, pos, n, P, D, oP, oD, oid);
    e = hipMemcpyAsync(&kept, pos + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = sync_stream(s, "reduce compaction");
  }
  std::vector<uint32_t> ids(kept);
  if (e == hipSuccess && kept) {
    e = hipMemcpyAsync(P, oP, sizeof(float4) * kept, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(D, oD, sizeof(float4) * kept, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ids.data(), oid, sizeof(uint32_t) * kept, hipMemcpyDeviceToHost, s);
  }
  hipEventRecord(ev[1], s);
  if (e == hipSuccess) e = sync_stream(s, "reduce");
  if (e == hipSuccess) e = hipGetLastError();
  release();
  if (e != hipSuccess) return fail(h, NF_E_HIP, std::string("reduce: ") + hipGetErrorString(e));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
  // the kept rods are the packing now (a subset of the old one: the length
  // and radius bounds of the handle still hold); every contact-derived
  // result refers to the old ids
  h->n_rods[h->post_inst] = kept;
  release_contacts(h);
  h->reduce_ids.swap(ids);
  if (info) {
    info->n_before = n;
    info->n_after = kept;
    info->rounds = rounds;
    info->pad = 0;
    info->device_ms = ms;
  }
  return NF_OK;
}

extern "C" int nf_get_reduce_kept(nf_handle* h, uint32_t* old_ids, uint64_t cap) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  if (cap < h->reduce_ids.size()) return fail(h, NF_E_CAPACITY, "buffer too small");
  if (old_ids && !h->reduce_ids.empty())
    std::copy(h->reduce_ids.begin(), h->reduce_ids.end(), old_ids);
  return NF_OK;
}
