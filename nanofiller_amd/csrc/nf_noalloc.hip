// nf_noalloc.hip - entry points that allocate no device memory.
//
// nf_histograms_noalloc is nf_histograms with its scratch (the bin-edge table
// and the 64-bit accumulators the kernels count into) carved from one
// file-scope device array instead of the handle's allocator, for callers that
// must not allocate: code under stream capture, tight serving loops,
// harnesses that forbid hipMalloc.  It calls no hipMalloc, hipFree or other
// allocating routine.
//
// Each GPU holds its own copy of the array, at its own address, for the life
// of the process.  Host code never passes the array's name to a kernel or a
// copy: the name is a host-side stand-in.  The device address is looked up
// once per device with hipGetSymbolAddress, with that device current, and
// kept per device.  Every call on one GPU shares that GPU's array: calls that
// overlap on one GPU (two host threads, or two handles' streams) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "nf_host.h"

using namespace nf;

namespace {

// Edge table, then the accumulators, each 256-byte aligned.
constexpr size_t NOALLOC_EDGES_OFF = 0;
constexpr size_t NOALLOC_OUT_OFF = ((HIST_EDGES_MAX * sizeof(double) + 255) / 256) * 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_OUT_OFF + HIST_OUT_MAX * sizeof(unsigned long long);
static_assert(NOALLOC_BYTES <= 64 * 1024, "scratch of the allocation-free entry points is 64 KiB");

constexpr int MAX_DEVICES = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[64 * 1024];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on device dev (looked up once, with dev
// current; the caller's current device is restored).
int scratch_on(nf_handle* h, int dev, unsigned char** out) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(h, NF_E_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
    void* ptr = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) {
      const hipError_t e2 = hipSetDevice(cur);
      if (e == hipSuccess) e = e2;
    }
    if (e != hipSuccess || !ptr)
      return fail(h, NF_E_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = (unsigned char*)g_scratch_ptr[dev];
  return NF_OK;
}

}  // namespace

extern "C" int nf_histograms_noalloc(nf_handle* h, const nf_hist_spec* spec, int64_t* out_centre,
                                     int64_t* out_polar, int64_t* out_mutual,
                                     int64_t* out_phitheta) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  h->err.clear();
  double edges[HIST_EDGES_MAX];
  HistPlan pl;
  int rc = hist_prepare(h, spec, out_mutual != nullptr, out_phitheta != nullptr, edges, &pl);
  if (rc != NF_OK) return rc;
  if (pl.n_edges > HIST_EDGES_MAX || pl.n_out > HIST_OUT_MAX)
    return fail(h, NF_E_INVALID, "histogram request exceeds the 64 KiB scratch");
  unsigned char* base = nullptr;
  rc = scratch_on(h, h->device, &base);
  if (rc != NF_OK) return rc;
  return hist_run(h, pl, edges, (double*)(base + NOALLOC_EDGES_OFF),
                  (unsigned long long*)(base + NOALLOC_OUT_OFF), out_centre, out_polar, out_mutual,
                  out_phitheta);
}
