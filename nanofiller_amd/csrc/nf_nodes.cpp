// nf_nodes.cpp - host orchestration of the node-level segment network (see
// nf_nodes.hip) on the contacts of the last nf_contacts call, and its
// getters.  Three device phases; a synchronisation after each of the first
// two sizes the next one's arrays, and after the third, the compact tables
// the getters copy (kept edges, electrode points) are built on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "nf_host.h"

using namespace nf;

namespace {

void free_nodes(nf_handle* h) {
  ContactBuffers& c = h->cb;
  void* ptrs[] = {c.nd_pflag, c.nd_edge_a, c.nd_edge_b, c.nd_esrc,    c.nd_esnk,
                  c.nd_src_node, c.nd_snk_node, c.nd_egs, c.nd_egk, c.nd_node_rod,
                  c.nd_node_t, c.nd_seg_a,  c.nd_seg_b, c.nd_seg_L, c.nd_seg_A,
                  c.nd_ce_a, c.nd_ce_b, c.nd_ce_k, c.nd_cel_node, c.nd_cel_gap, c.nd_cel_side};
  for (void* p : ptrs) dfree(h, p);
  c.nd_pflag = c.nd_edge_a = c.nd_edge_b = c.nd_esrc = c.nd_esnk = nullptr;
  c.nd_src_node = c.nd_snk_node = c.nd_node_rod = c.nd_seg_a = c.nd_seg_b = nullptr;
  c.nd_egs = c.nd_egk = c.nd_node_t = c.nd_seg_L = c.nd_seg_A = nullptr;
  c.nd_ce_a = c.nd_ce_b = c.nd_ce_k = c.nd_cel_node = nullptr;
  c.nd_cel_gap = nullptr;
  c.nd_cel_side = nullptr;
  c.nd_n_edges = c.nd_n_elec = 0;
  c.nd_nodes = c.nd_segments = c.nd_points = 0;
  c.nd_valid = false;
}

}  // namespace

extern "C" int nf_node_network(nf_handle* h, nf_node_info* info) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  ContactBuffers& c = h->cb;
  if (!c.valid) return fail(h, NF_E_STATE, "no contacts computed");
  const nf_params& p = h->params[h->post_inst];
  if (p.elec_axis < 0 || p.elec_axis > 2) return fail(h, NF_E_INVALID, "elec_axis must be 0..2");
  free_nodes(h);
  // the screening walk runs over the centre cell list of the contact rods
  int rc0 = build_contact_cells(h);
  if (rc0 != NF_OK) return rc0;
  const uint32_t n = c.n_rods;
  const uint64_t nc = c.n_contacts;
  hipStream_t s = h->stream;
  NodeArgs a{};
  a.n = n;
  a.rowptr = c.rowptr;
  a.J = c.J;
  a.img = c.img;
  a.P = h->gb.bP + h->desc[h->post_inst].rod_off;
  a.D = h->gb.bD + h->desc[h->post_inst].rod_off;
  a.axis = p.elec_axis;
  for (int d = 0; d < 3; ++d) {
    a.box[d] = p.box[d];
    a.ncell[d] = c.ncell[d];
    a.cell_scale[d] = (double)c.ncell[d] / p.box[d];
  }
  a.L = p.box[p.elec_axis];
  a.gcut = p.contact_gap;
  // the bounds the contact pass used: from the data when the rods came from
  // nf_set_rods or a state file, so longer or thicker rods are reached too
  a.hl_max = (double)c.hl_max;
  a.a_max = (double)c.a_max;
  a.cP = c.cP;
  a.cD = c.cD;
  a.cid = c.cid;
  a.cstart = c.cstart;
  a.ccount = c.ccount;
  uint32_t *pstart = nullptr, *pfill = nullptr, *npt = nullptr, *ngrp = nullptr, *nseg = nullptr,
           *gstart = nullptr, *sstart = nullptr, *tmp = nullptr, *pt_grp = nullptr, *rowof = nullptr;
  unsigned long long* cnt = nullptr;
  double *ptI = nullptr, *ptJ = nullptr, *pt_t = nullptr;
  unsigned long long* pt_key = nullptr;
  auto release = [&]() {
    void* ptrs[] = {pstart, pfill, npt, ngrp, nseg, gstart, sstart, tmp, pt_grp, ptI, ptJ, pt_t, pt_key, cnt, rowof};
    for (void* q : ptrs) dfree(h, q);
  };
  const uint64_t n1 = (uint64_t)n + 1, ncs = std::max<uint64_t>(nc, 1);
  int rc = NF_OK;
#define ALLOC(ptr, cnt) \
  if (rc == NF_OK) rc = dalloc(h, &(ptr), (size_t)std::max<uint64_t>((cnt), 1));
  ALLOC(c.nd_pflag, ncs);
  ALLOC(ptI, ncs);
  ALLOC(ptJ, ncs);
  ALLOC(rowof, ncs);
  ALLOC(c.nd_edge_a, ncs);
  ALLOC(c.nd_edge_b, ncs);
  ALLOC(npt, n1);
  ALLOC(pstart, n1);
  ALLOC(pfill, n1);
  ALLOC(ngrp, n1);
  ALLOC(nseg, n1);
  ALLOC(gstart, n1);
  ALLOC(sstart, n1);
  ALLOC(c.nd_esrc, n1);
  ALLOC(c.nd_esnk, n1);
  ALLOC(c.nd_src_node, n1);
  ALLOC(c.nd_snk_node, n1);
  ALLOC(c.nd_egs, n1);
  ALLOC(c.nd_egk, n1);
  ALLOC(tmp, scan_tmp_words(std::max<uint64_t>(n1, ncs)));  // scans over rods and over contacts
  ALLOC(cnt, 4);
  auto fail_out = [&](int code, const std::string& m) {
    release();
    free_nodes(h);
    return fail(h, code, m);
  };
  if (rc != NF_OK) return fail_out(rc, "node network: out of device memory");
  a.pflag = c.nd_pflag;
  a.ptI = ptI;
  a.rowof = rowof;
  a.ptJ = ptJ;
  a.edge_a = c.nd_edge_a;
  a.edge_b = c.nd_edge_b;
  a.npt = npt;
  a.pstart = pstart;
  a.pfill = pfill;
  a.ngrp = ngrp;
  a.nseg = nseg;
  a.gstart = gstart;
  a.sstart = sstart;
  a.esrc = c.nd_esrc;
  a.esnk = c.nd_esnk;
  a.src_node = c.nd_src_node;
  a.snk_node = c.nd_snk_node;
  a.egs = c.nd_egs;
  a.egk = c.nd_egk;
  hipEventRecord(h->ev[0], s);
  fill_u32(npt, 0u, n1, s);
  fill_u32(pfill, 0u, n1, s);
  fill_u32(ngrp, 0u, n1, s);
  fill_u32(nseg, 0u, n1, s);
  if (n > 0) launch_node_points(a, rowof, s);
  scan_u32(npt, pstart, n1, tmp, s);
  uint32_t npts = 0;
  if (hipMemcpyAsync(&npts, pstart + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      sync_stream(s, "node points") != hipSuccess)
    return fail_out(NF_E_HIP, "node network: point phase failed");
  ALLOC(pt_t, npts);
  ALLOC(pt_key, npts);
  ALLOC(pt_grp, npts);
  if (rc != NF_OK) return fail_out(rc, "node network: out of device memory");
  a.pt_t = pt_t;
  a.pt_key = pt_key;
  a.pt_grp = pt_grp;
  if (n > 0) launch_node_lists(a, s);
  scan_u32(ngrp, gstart, n1, tmp, s);
  scan_u32(nseg, sstart, n1, tmp, s);
  uint32_t tot[2] = {0, 0};
  if (hipMemcpyAsync(&tot[0], gstart + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&tot[1], sstart + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      sync_stream(s, "node lists") != hipSuccess)
    return fail_out(NF_E_HIP, "node network: list phase failed");
  ALLOC(c.nd_node_rod, tot[0]);
  ALLOC(c.nd_node_t, tot[0]);
  ALLOC(c.nd_seg_a, tot[1]);
  ALLOC(c.nd_seg_b, tot[1]);
  ALLOC(c.nd_seg_L, tot[1]);
  ALLOC(c.nd_seg_A, tot[1]);
#undef ALLOC
  if (rc != NF_OK) return fail_out(rc, "node network: out of device memory");
  a.node_rod = c.nd_node_rod;
  a.node_t = c.nd_node_t;
  a.seg_a = c.nd_seg_a;
  a.seg_b = c.nd_seg_b;
  a.seg_L = c.nd_seg_L;
  a.seg_A = c.nd_seg_A;
  if (n > 0) launch_node_nodes(a, s);
  fill_u32((uint32_t*)cnt, 0u, 8, s);
  if (n > 0) launch_node_count(a, (uint32_t)nc, cnt, s);
  unsigned long long hc[4] = {0, 0, 0, 0};
  if (hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s) != hipSuccess ||
      sync_stream(s, "node network") != hipSuccess)
    return fail_out(NF_E_HIP, "node network: node phase failed");
  // compact tables: kept edges (contact order), electrode points (source
  // then sink, rod order)
  uint32_t *kept = nullptr, *rkept = nullptr, *rsrc = nullptr, *rsnk = nullptr;
  const uint64_t n_elec = hc[2] + hc[3];
  if (rc == NF_OK) rc = dalloc(h, &kept, ncs);
  if (rc == NF_OK) rc = dalloc(h, &rkept, ncs + 1);
  if (rc == NF_OK) rc = dalloc(h, &rsrc, n1);
  if (rc == NF_OK) rc = dalloc(h, &rsnk, n1);
  if (rc == NF_OK) rc = dalloc(h, &c.nd_ce_a, std::max<uint64_t>(hc[0], 1));
  if (rc == NF_OK) rc = dalloc(h, &c.nd_ce_b, std::max<uint64_t>(hc[0], 1));
  if (rc == NF_OK) rc = dalloc(h, &c.nd_ce_k, std::max<uint64_t>(hc[0], 1));
  if (rc == NF_OK) rc = dalloc(h, &c.nd_cel_node, std::max<uint64_t>(n_elec, 1));
  if (rc == NF_OK) rc = dalloc(h, &c.nd_cel_gap, std::max<uint64_t>(n_elec, 1));
  if (rc == NF_OK) rc = dalloc(h, &c.nd_cel_side, std::max<uint64_t>(n_elec, 1));
  auto free_ranks = [&]() {
    dfree(h, kept);
    dfree(h, rkept);
    dfree(h, rsrc);
    dfree(h, rsnk);
  };
  if (rc != NF_OK) {
    free_ranks();
    return fail_out(rc, "node network: out of device memory");
  }
  launch_node_kept_flag(c.nd_pflag, (uint32_t)nc, kept, s);
  scan_u32(kept, rkept, nc, tmp, s);
  scan_u32(c.nd_esrc, rsrc, n, tmp, s);
  scan_u32(c.nd_esnk, rsnk, n, tmp, s);
  launch_node_tables(a, (uint32_t)nc, kept, rkept, c.nd_ce_a, c.nd_ce_b, c.nd_ce_k, rsrc, rsnk, (uint32_t)hc[2],
                     c.nd_cel_node, c.nd_cel_gap, c.nd_cel_side, s);
  hipEventRecord(h->ev[1], s);
  if (hipGetLastError() != hipSuccess || sync_stream(s, "node tables") != hipSuccess) {
    free_ranks();
    return fail_out(NF_E_HIP, "node network: table phase failed");
  }
  free_ranks();
  c.nd_n_edges = hc[0];
  c.nd_n_elec = n_elec;
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
  release();
  c.nd_points = npts;
  c.nd_nodes = tot[0];
  c.nd_segments = tot[1];
  c.nd_valid = true;
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->n_points = npts;
    info->n_nodes = tot[0];
    info->n_segments = tot[1];
    info->n_edges = (uint32_t)hc[0];
    info->n_screened = (uint32_t)hc[1];
    info->n_source = (uint32_t)hc[2];
    info->n_sink = (uint32_t)hc[3];
    info->device_ms = ms;
  }
  return NF_OK;
}

extern "C" int nf_get_nodes(nf_handle* h, uint32_t* rod, double* t, uint64_t cap) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  const ContactBuffers& c = h->cb;
  if (!c.nd_valid) return fail(h, NF_E_STATE, "no node network built");
  if (cap < c.nd_nodes) return fail(h, NF_E_CAPACITY, "node buffer too small");
  if (c.nd_nodes == 0) return NF_OK;
  if (rod) HIPCHK(h, hipMemcpy(rod, c.nd_node_rod, sizeof(uint32_t) * c.nd_nodes, hipMemcpyDeviceToHost));
  if (t) HIPCHK(h, hipMemcpy(t, c.nd_node_t, sizeof(double) * c.nd_nodes, hipMemcpyDeviceToHost));
  return NF_OK;
}

extern "C" int nf_get_segments(nf_handle* h, uint32_t* na, uint32_t* nb, double* L, double* A,
                               uint64_t cap) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  const ContactBuffers& c = h->cb;
  if (!c.nd_valid) return fail(h, NF_E_STATE, "no node network built");
  const uint64_t m = c.nd_segments;
  if (cap < m) return fail(h, NF_E_CAPACITY, "segment buffer too small");
  if (m == 0) return NF_OK;
  if (na) HIPCHK(h, hipMemcpy(na, c.nd_seg_a, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  if (nb) HIPCHK(h, hipMemcpy(nb, c.nd_seg_b, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  if (L) HIPCHK(h, hipMemcpy(L, c.nd_seg_L, sizeof(double) * m, hipMemcpyDeviceToHost));
  if (A) HIPCHK(h, hipMemcpy(A, c.nd_seg_A, sizeof(double) * m, hipMemcpyDeviceToHost));
  return NF_OK;
}

extern "C" int nf_get_node_edges(nf_handle* h, uint32_t* na, uint32_t* nb, uint32_t* contact,
                                 uint64_t cap, uint64_t* count) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  const ContactBuffers& c = h->cb;
  if (!c.nd_valid) return fail(h, NF_E_STATE, "no node network built");
  const uint64_t m = c.nd_n_edges;
  if (count) *count = m;
  if (!na && !nb && !contact) return NF_OK;
  if (cap < m) return fail(h, NF_E_CAPACITY, "edge buffer too small");
  if (m == 0) return NF_OK;
  if (na) HIPCHK(h, hipMemcpy(na, c.nd_ce_a, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  if (nb) HIPCHK(h, hipMemcpy(nb, c.nd_ce_b, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  if (contact) HIPCHK(h, hipMemcpy(contact, c.nd_ce_k, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  return NF_OK;
}

extern "C" int nf_get_node_electrodes(nf_handle* h, uint32_t* node, double* gap, int32_t* side,
                                      uint64_t cap, uint64_t* count) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  const ContactBuffers& c = h->cb;
  if (!c.nd_valid) return fail(h, NF_E_STATE, "no node network built");
  const uint64_t m = c.nd_n_elec;
  if (count) *count = m;
  if (!node && !gap && !side) return NF_OK;
  if (cap < m) return fail(h, NF_E_CAPACITY, "electrode buffer too small");
  if (m == 0) return NF_OK;
  if (node) HIPCHK(h, hipMemcpy(node, c.nd_cel_node, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  if (gap) HIPCHK(h, hipMemcpy(gap, c.nd_cel_gap, sizeof(double) * m, hipMemcpyDeviceToHost));
  if (side) HIPCHK(h, hipMemcpy(side, c.nd_cel_side, sizeof(int32_t) * m, hipMemcpyDeviceToHost));
  return NF_OK;
}
