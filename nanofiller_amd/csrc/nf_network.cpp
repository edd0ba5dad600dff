// nf_network.cpp - host orchestration of network extraction on the contacts
// of the last nf_contacts call (instance 0).
//
// Replaces collect_cunduct_data (src/conductivity.cu:140-423), whose output
// is a set of COO contact files for an external solver, and the component
// labelling of adjacency_graph (src/adjacency_graph.cu:57-219).  Two device
// phases with one synchronisation between them to size the matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "nf_host.h"

using namespace nf;

extern "C" int nf_network(nf_handle* h, nf_network_info* info) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  ContactBuffers& c = h->cb;
  if (!c.valid) return fail(h, NF_E_STATE, "no contacts computed");
  const nf_params& p = h->params[h->post_inst];
  const int ax = p.elec_axis;
  if (ax < 0 || ax > 2) return fail(h, NF_E_INVALID, "elec_axis must be 0..2");
  const uint32_t n = c.n_rods;
  const uint64_t nc = c.n_contacts;
  if ((uint64_t)n + 3 > 0xFFFFFFFFull) return fail(h, NF_E_INVALID, "too many rods for the network");
  hipStream_t s = h->stream;
  // results of an earlier call
  dfree(h, c.labels);
  dfree(h, c.nrow);
  dfree(h, c.ncol);
  dfree(h, c.nval);
  c.labels = c.nrow = c.ncol = nullptr;
  c.nval = nullptr;
  c.net_valid = false;

  NetArgs a{};
  a.rowptr = c.rowptr;
  a.J = c.J;
  a.img = c.img;
  a.gap64 = c.gap64;
  a.P = h->gb.bP + h->desc[h->post_inst].rod_off;
  a.D = h->gb.bD + h->desc[h->post_inst].rod_off;
  a.n = n;
  a.axis = ax;
  a.L = p.box[ax];
  a.gcut = p.contact_gap;
  a.lam = p.tunnel_lambda;
  a.g0 = p.tunnel_g0;
  const uint64_t n1 = (uint64_t)n + 1;
  int rc = NF_OK;
#define ALLOC(ptr, cnt) \
  if (rc == NF_OK) rc = dalloc(h, &(ptr), (size_t)(cnt));
  ALLOC(c.labels, std::max<uint64_t>(n, 1));
  ALLOC(c.nrow, (uint64_t)n + 3);
  ALLOC(a.lowcnt, n1);
  ALLOC(a.lowstart, n1);
  ALLOC(a.lowfill, n1);
  ALLOC(a.lcol, std::max<uint64_t>(nc, 1));
  ALLOC(a.limg, std::max<uint64_t>(nc, 1));
  ALLOC(a.lg, std::max<uint64_t>(nc, 1));
  ALLOC(a.fsrc, n1);
  ALLOC(a.fsnk, n1);
  ALLOC(a.psrc, n1);
  ALLOC(a.psnk, n1);
  ALLOC(a.gsrc, n1);
  ALLOC(a.gsnk, n1);
  ALLOC(a.csize, n1);
  ALLOC(a.cflag, n1);
  ALLOC(a.rowcnt, (uint64_t)n + 3);
  ALLOC(a.summary, NET_SUMMARY);
  ALLOC(a.tmp, scan_tmp_words((uint64_t)n + 2));
  ALLOC(a.cg, n1);
  ALLOC(a.bsum, n1 / 1024 + 2);
#undef ALLOC
  auto release_work = [&]() {
    void* w[] = {a.lowcnt, a.lowstart, a.lowfill, a.lcol, a.limg, a.lg, a.fsrc, a.fsnk,
                 a.psrc,   a.psnk,     a.gsrc,    a.gsnk, a.csize, a.cflag, a.rowcnt,
                 a.summary, a.tmp,     a.cg,      a.bsum};
    for (void* q : w) dfree(h, q);
  };
  if (rc != NF_OK) {
    release_work();
    return rc;
  }
  a.par = c.labels;
  a.nrow = c.nrow;
  hipEventRecord(h->ev[0], s);
  launch_network_build(a, s);
  unsigned long long sum[NET_SUMMARY] = {};
  hipError_t e = hipMemcpyAsync(sum, a.summary, sizeof(sum), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = sync_stream(s, "network build");
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) {
    release_work();
    return fail(h, NF_E_HIP, std::string("network build: ") + hipGetErrorString(e));
  }
  const uint64_t nnz = sum[SUM_NNZ];
  trace("network: %u rods, %llu edges, %llu components, nnz %llu", n, sum[SUM_EDGES],
        sum[SUM_COMPONENTS], (unsigned long long)nnz);
  rc = dalloc(h, &c.ncol, nnz);
  if (rc == NF_OK) rc = dalloc(h, &c.nval, nnz);
  if (rc != NF_OK) {
    release_work();
    return rc;
  }
  a.ncol = c.ncol;
  a.nval = c.nval;
  launch_network_fill(a, (uint32_t)sum[SUM_SOURCE], (uint32_t)sum[SUM_SINK], s);
  hipEventRecord(h->ev[1], s);
  e = sync_stream(s, "network fill");
  if (e == hipSuccess) e = hipGetLastError();
  release_work();
  if (e != hipSuccess) return fail(h, NF_E_HIP, std::string("network fill: ") + hipGetErrorString(e));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
  h->net_ms = ms;
  nf_network_info& ni = h->net_info;
  ni = nf_network_info{};
  ni.n_nodes = n + 2;
  ni.n_components = (uint32_t)sum[SUM_COMPONENTS];
  ni.largest = (uint32_t)sum[SUM_LARGEST];
  ni.percolating = (uint32_t)sum[SUM_PERCOLATING];
  ni.nnz = nnz;
  ni.n_edges = sum[SUM_EDGES];
  ni.n_source = (uint32_t)sum[SUM_SOURCE];
  ni.n_sink = (uint32_t)sum[SUM_SINK];
  c.nnz = nnz;
  c.net_valid = true;
  if (info) *info = ni;
  return NF_OK;
}

extern "C" int nf_get_network(nf_handle* h, uint32_t* labels, uint32_t* row_ptr, uint32_t* col,
                              double* val, uint64_t cap) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  ContactBuffers& c = h->cb;
  if (!c.net_valid) return fail(h, NF_E_STATE, "no network computed");
  if ((col || val) && cap < c.nnz) return fail(h, NF_E_CAPACITY, "network buffer too small");
  const uint32_t n = c.n_rods;
  if (labels && n)
    HIPCHK(h, hipMemcpy(labels, c.labels, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
  if (row_ptr)
    HIPCHK(h, hipMemcpy(row_ptr, c.nrow, sizeof(uint32_t) * ((size_t)n + 3), hipMemcpyDeviceToHost));
  if (c.nnz) {
    if (col) HIPCHK(h, hipMemcpy(col, c.ncol, sizeof(uint32_t) * c.nnz, hipMemcpyDeviceToHost));
    if (val) HIPCHK(h, hipMemcpy(val, c.nval, sizeof(double) * c.nnz, hipMemcpyDeviceToHost));
  }
  return NF_OK;
}
