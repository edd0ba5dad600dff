// nf_host.h - host-side handle layout and helpers shared by the library's
// C++ translation units (nf_api.cpp, nf_contacts.cpp, nf_io.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "nf_internal.h"

// Options of nf_set_option() (declared in nanofiller.h).
struct nf_handle {
  std::vector<nf_params> params;
  std::vector<nf::InstDesc> desc;  // host copy of the device descriptors
  nf::GenBuffers gb;
  nf::ContactBuffers cb;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_prec = -1, graph_diag = -1;
  hipEvent_t ev[9] = {};
  std::string err;
  int opt_graph = 0, opt_profile = 0, opt_diag = 0, opt_contact_prec = -1;
  int opt_contact_engine = NF_ENGINE_FOOTPRINT;
  int opt_footprint_fine = 0;  // footprint fine cell edge in 1/1000 of the typical rod length (0: 200)
  int opt_footprint_occ = 0;   // footprint home cell: rods crossing it at the mean density (0: 64)
  int post_inst = 0;  // instance that contacts, statistics, network, reduce, writers, state files and expand use
  int device = 0;
  std::vector<uint32_t> n_rods;  // current rod count per instance
  std::vector<uint32_t> planned_cap;  // cell slots per instance as planned (NF_OPT_SLOT_SCALE scales them)
  std::vector<double> milestones;  // volume fractions (percent) of the state dumps, ascending
  std::string milestone_dir;       // where nf_generate writes them
  bool generated = false;
  // device memory: live blocks (pointer -> size-class bytes) and a pool of
  // freed blocks by size class, reused by later calls of the same handle
  std::unordered_map<void*, size_t> allocs;
  std::multimap<size_t, void*> pool;
  size_t pool_bytes = 0;
  unsigned long long cb_stats[6] = {0, 0, 0, 0, 0, 0};  // contact pair tests, rechecks, max dd bits
  float cb_ms[4] = {0, 0, 0, 0};                  // contact kernel times: bin, pair pass, CSR build, row sort
  uint64_t contact_hint = 0;                      // contacts found by the last call (append capacity)
  nf_network_info net_info = {};
  std::vector<uint32_t> reduce_ids;  // old id of every rod kept by the last nf_reduce
  float net_ms = 0.0f;  // device time of the last nf_network call
};

extern std::string g_last_error;

namespace nf {

int fail(nf_handle* h, int code, const std::string& msg);
// Waits for the stream with a wall-clock limit (NF_SYNC_TIMEOUT_S, default
// 120 s); prints progress when NF_TRACE is set.
hipError_t sync_stream(hipStream_t s, const char* what);
void trace(const char* fmt, ...);
int dalloc_bytes(nf_handle* h, void** p, size_t bytes);
void dfree(nf_handle* h, void* p);
void release_contacts(nf_handle* h);
// Centre cell list of the contact rods (edge about one pair reach, sorted
// stencil): built by the cell engine of nf_contacts, or on demand by the
// node network.
int build_contact_cells(nf_handle* h);
// Largest half length and radius of instance 0's n rods at P, D: the
// generation's bounds after nf_generate, else scanned (one host copy).
int contact_bounds(nf_handle* h, uint32_t n, const float4* P, const float4* D, float* hlmax, float* amax);
// Default cell edge for a typical pair reach and a mean rod density.
double default_cell_edge(double reach_typ, double density, double factor);
// Neighbour offsets of a home block of bx cells along x (offsets from its
// first cell) whose cells come within `reach` of the block, sorted by that
// minimum distance (its square, slightly low, in .w as float bits); `half`
// (bx = 1 only) keeps one of each +-offset pair.
std::vector<int4> build_stencil(const int32_t nc[3], const double box[3], double reach, bool half, int bx = 1);

// Histogram limits (nf_hist_spec): 1-D bins, phi x theta cells, and the
// bin-edge table: polar and mutual edges (bins + 1 each), then the theta
// (theta_bins + 1) and phi (phi_bins / 2 + 1) edges of the 2-D histogram.
constexpr int HIST_BINS_MAX = 256;
constexpr int HIST_2D_MAX = 4096;
constexpr size_t HIST_EDGES_MAX = 2 * (HIST_BINS_MAX + 1) + HIST_2D_MAX / 2 + 3;
constexpr size_t HIST_OUT_MAX = 3 * HIST_BINS_MAX + HIST_2D_MAX;

struct HistPlan {
  int nb, npb, ntb, axis;  // 1-D bins, 2-D phi/theta bins (0: no 2-D), centre axis
  bool mutual;             // mutual-angle histogram over the last contact list
  double mu[3];            // unit polar axis
  size_t n_edges, n_out;   // edge-table and output lengths (doubles, u64 words)
};
// Validates a histogram request and fills the edge table (edges[HIST_EDGES_MAX]).
int hist_prepare(nf_handle* h, const nf_hist_spec* spec, bool want_mutual, bool want_2d,
                 double* edges, HistPlan* plan);
// Uploads the edges to dE[n_edges], clears dOut[n_out], counts, and copies the
// counts to the caller's host arrays (any may be NULL); waits for the stream.
int hist_run(nf_handle* h, const HistPlan& pl, const double* edges, double* dE,
             unsigned long long* dOut, int64_t* out_centre, int64_t* out_polar, int64_t* out_mutual,
             int64_t* out_phitheta);

template <class T>
int dalloc(nf_handle* h, T** p, size_t count) {
  void* q = nullptr;
  const int rc = dalloc_bytes(h, &q, count * sizeof(T));
  *p = (T*)q;
  return rc;
}

}  // namespace nf

#define HIPCHK(h, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return nf::fail((h), NF_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)
