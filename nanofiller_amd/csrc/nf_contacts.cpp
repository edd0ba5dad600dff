// nf_contacts.cpp - host orchestration of contact detection and histograms
// on the packing of instance 0.
//
// Contacts: bin -> one pair pass that appends every contact to an unsorted
// list -> CSR over I (count, scan, scatter) -> canonical row sort -> contact
// areas.  This replaces the COUNT and RECORD passes of get_contact_counts /
// get_contact_arrays (src/get_contact_arrays.cu:12-140); a pass whose list
// overflows the append buffer is repeated once at the exact size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nf_host.h"

using namespace nf;

namespace nf {

// Reach bounds of the contact rods: the instance's bounds after generation,
// else the largest half length and radius of the rods (copied once).
int contact_bounds(nf_handle* h, uint32_t n, const float4* P, const float4* D, float* hlmax, float* amax) {
  *hlmax = h->desc[h->post_inst].hl_max;
  *amax = (float)h->params[h->post_inst].radius;
  if (h->generated) return NF_OK;
  std::vector<float4> hP(n), hD(n);
  if (n) {
    HIPCHK(h, hipMemcpy(hP.data(), P, sizeof(float4) * n, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(hD.data(), D, sizeof(float4) * n, hipMemcpyDeviceToHost));
  }
  *hlmax = 0.0f;
  *amax = 0.0f;
  for (uint32_t i = 0; i < n; ++i) {
    *hlmax = std::max(*hlmax, hP[i].w);
    *amax = std::max(*amax, hD[i].w);
  }
  return NF_OK;
}

static double contact_reach(const ContactBuffers& c) {
  const double Rmax = ((double)c.a_max + (double)c.a_max) + c.gap_cut;
  return ((double)c.hl_max + (double)c.hl_max) * REACH_PAD + Rmax;
}

// Centre cell list of the contact rods on cells of about `edge`, with the
// stencil of every cell within `reach` of a home block of bx cells along x
// (half, bx = 1: each cell pair once).
// Scratch of bin_rods, given back by the caller once the stream has
// synchronized.
struct BinScratch {
  uint32_t* M = nullptr;
  uint32_t* tmp = nullptr;
  float4* tP = nullptr;
  float4* tD = nullptr;
  uint32_t* tid = nullptr;
};

static void free_bin_scratch(nf_handle* h, BinScratch* b) {
  dfree(h, b->M);
  dfree(h, b->tmp);
  dfree(h, b->tP);
  dfree(h, b->tD);
  dfree(h, b->tid);
  *b = BinScratch{};
}

// Cell list of n rods on the stream (launch_bin_rods): rods in cell order
// and each cell's start (ncells + 1), count and largest half length.
static int bin_rods(nf_handle* h, const float4* P, const float4* D, uint32_t n, const int32_t nc[3],
                    const double scale[3], float4* cP, float4* cD, uint32_t* cid, uint32_t* cstart,
                    uint32_t* ccount, float* chl, BinScratch* b) {
  const uint64_t ncells = (uint64_t)nc[0] * nc[1] * nc[2];
  const BinPlan bp = bin_plan(n, ncells);
  const uint64_t n1 = std::max<uint64_t>(n, 1);
  int rc = NF_OK;
  if (rc == NF_OK) rc = dalloc(h, &b->M, bp.m_words + 1);
  if (rc == NF_OK) rc = dalloc(h, &b->tmp, scan_tmp_words(bp.m_words));
  if (rc == NF_OK) rc = dalloc(h, &b->tP, n1);
  if (rc == NF_OK) rc = dalloc(h, &b->tD, n1);
  if (rc == NF_OK) rc = dalloc(h, &b->tid, n1);
  if (rc != NF_OK) return rc;
  launch_bin_rods(bp, P, D, n, nc, scale, b->M, b->tmp, b->tP, b->tD, b->tid, cP, cD, cid, cstart, ccount, chl,
                  h->stream);
  return NF_OK;
}

static int build_cell_list(nf_handle* h, double edge, double reach, bool half, int bx = 1) {
  ContactBuffers& c = h->cb;
  const nf_params& p = h->params[h->post_inst];
  const uint32_t n = c.n_rods;
  const float4* P = h->gb.bP + h->desc[h->post_inst].rod_off;
  const float4* D = h->gb.bD + h->desc[h->post_inst].rod_off;
  hipStream_t s = h->stream;
  uint64_t ncells = 1;
  double scale[3];
  for (int a = 0; a < 3; ++a) {
    double nn = std::floor(p.box[a] / edge);
    if (nn < 1.0) nn = 1.0;
    if (nn > 1024.0) nn = 1024.0;
    c.ncell[a] = (int32_t)nn;
    scale[a] = (double)c.ncell[a] / p.box[a];
    c.cscale[a] = scale[a];
    ncells *= (uint64_t)c.ncell[a];
  }
  if (ncells > (1ull << 26)) return fail(h, NF_E_INVALID, "contact grid too large");
  c.ncells = (uint32_t)ncells;
  std::vector<int4> st = build_stencil(c.ncell, p.box, reach, half, bx);
  c.st_n = (uint32_t)st.size();
  int rc = NF_OK;
#define ALLOC(ptr, cnt) \
  if (rc == NF_OK) rc = dalloc(h, &(ptr), (cnt));
  ALLOC(c.cP, std::max<uint64_t>(n, 1));
  ALLOC(c.cD, std::max<uint64_t>(n, 1));
  ALLOC(c.cid, std::max<uint64_t>(n, 1));
  ALLOC(c.cstart, ncells + 1);
  ALLOC(c.ccount, ncells);
  ALLOC(c.chl, ncells);
  ALLOC(c.stencil, st.size());
#undef ALLOC
  if (rc != NF_OK) return rc;
  HIPCHK(h, hipMemcpyAsync(c.stencil, st.data(), sizeof(int4) * st.size(), hipMemcpyHostToDevice, s));
  BinScratch bs;
  rc = bin_rods(h, P, D, n, c.ncell, scale, c.cP, c.cD, c.cid, c.cstart, c.ccount, c.chl, &bs);
  if (rc != NF_OK) {
    free_bin_scratch(h, &bs);
    return rc;
  }
  const hipError_t e = sync_stream(s, "contact cell list");
  free_bin_scratch(h, &bs);
  if (e != hipSuccess) return fail(h, NF_E_HIP, std::string("contact cell list: ") + hipGetErrorString(e));
  return NF_OK;
}

// Cell list of the cell-pair engine: cells of about one typical pair reach,
// half stencil of the longest pair reach.
int build_contact_cells(nf_handle* h) {
  ContactBuffers& c = h->cb;
  if (c.cP && c.cell_list == 1) return NF_OK;
  if (c.cP) {
    // the footprint engine's list (home-cell grid, run stencil): rebuild
    void* ptrs[] = {c.cP, c.cD, c.cid, c.cstart, c.ccount, c.chl, c.cmask, c.stencil};
    for (void* q : ptrs) dfree(h, q);
    c.cP = c.cD = nullptr;
    c.cid = c.cstart = c.ccount = c.cmask = nullptr;
    c.chl = nullptr;
    c.stencil = nullptr;
    c.cell_list = 0;
  }
  const nf_params& p = h->params[h->post_inst];
  const double Rmax = ((double)c.a_max + (double)c.a_max) + c.gap_cut;
  const double reach = contact_reach(c);
  // cell edge about one typical pair reach; the sorted stencil reaches the
  // longest rods and per-cell length bounds prune the rest
  double edge = (p.len_law == NF_LEN_LOGNORMAL ? p.len : (double)c.hl_max * 2.0) * REACH_PAD + Rmax;
  if (edge > reach) edge = reach;
  edge = default_cell_edge(edge, (double)c.n_rods / (p.box[0] * p.box[1] * p.box[2]), 1.0);
  const int rc = build_cell_list(h, edge, reach, true);
  if (rc == NF_OK) c.cell_list = 1;
  return rc;
}

// Footprint engine: home cells sized so that about 64 rods (the occupancy
// option) cross one at the mean density (rho (E^3 + 1.5 L E^2) rods of
// length L cross a cube of edge E, by Cauchy's mean projected area); a work
// unit that overflows its survivor table in a denser region is redone one
// cell at a time.  Fine cells of about a fifth of a rod length (the fine
// option sets the fraction) inside them.  Measured on 4M uniform 1 um rods,
// 200k 4 um rods, 500k log-normal rods in agglomerates and 8M rods in 16
// agglomerates: occupancy 32-64 and fine fraction 0.2 are within 20% of
// the best for each; from 96 on, units are redone cell by cell too often.
static void footprint_cells(nf_handle* h, uint32_t n, double* edge, int* fine) {
  const nf_params& p = h->params[h->post_inst];
  const ContactBuffers& c = h->cb;
  const double typ = p.len_law == NF_LEN_LOGNORMAL ? p.len : (double)c.hl_max * 2.0;
  const double V = p.box[0] * p.box[1] * p.box[2];
  const double rho = (double)n / V;
  const double occ = h->opt_footprint_occ > 0 ? (double)h->opt_footprint_occ : 64.0;
  double lo = 1e-3, hi = std::max(p.box[0], std::max(p.box[1], p.box[2]));
  for (int it = 0; it < 60; ++it) {
    const double e = 0.5 * (lo + hi);
    if (rho * (e * e * e + 1.5 * typ * e * e) > occ)
      hi = e;
    else
      lo = e;
  }
  *edge = std::max(lo, 4.0 * ((double)c.a_max + c.gap_cut));
  const double ff = h->opt_footprint_fine > 0 ? h->opt_footprint_fine * 1e-3 : 0.2;
  int f = (int)std::lround(*edge / std::max(typ * ff, 1e-6));
  *fine = f < 1 ? 1 : (f > 4 ? 4 : f);
}

// One contact pass with the given engine.  Returns 1 (not an NF_ status)
// when the footprint engine found a home cell with too many survivors.
static int contacts_pass(nf_handle* h, double gap, uint64_t* n_out, int prec, bool cells) {
  ContactBuffers& c = h->cb;
  const nf_params& p = h->params[h->post_inst];
  const InstDesc& d = h->desc[h->post_inst];
  const uint32_t n = h->n_rods[h->post_inst];
  const float4* P = h->gb.bP + d.rod_off;
  const float4* D = h->gb.bD + d.rod_off;
  hipStream_t s = h->stream;
  // release previous results
  release_contacts(h);
  c.n_rods = n;
  c.gap_cut = gap;
  float hlmax = 0.0f, amax = 0.0f;
  int rc = contact_bounds(h, n, P, D, &hlmax, &amax);
  if (rc != NF_OK) return rc;
  c.hl_max = hlmax;
  c.a_max = amax;
  const double Rmax = ((double)amax + (double)amax) + gap;
  const double reach = contact_reach(c);
  const double Lmax = std::max(p.box[0], std::max(p.box[1], p.box[2]));
  const double ew = 5.9604644775390625e-8 * (Lmax + reach) * 1.01;
  unsigned long long* ctr = nullptr;
  uint32_t* tmp = nullptr;
  uint32_t* queues = nullptr;
#define ALLOC(ptr, cnt) \
  if (rc == NF_OK) rc = dalloc(h, &(ptr), (cnt));
  ALLOC(c.cnt, (uint64_t)n + 1);
  ALLOC(c.rowptr, (uint64_t)n + 1);
  ALLOC(c.fill, std::max<uint64_t>(n, 1));
  ALLOC(ctr, 8);  // pair tests, rechecks, max dd, tile bytes, append count, work counter, overflow
  ALLOC(tmp, scan_tmp_words((uint64_t)n + 1));
  ALLOC(queues, 2 * QWORDS);
#undef ALLOC
  auto free_scratch = [&]() {
    dfree(h, ctr);
    dfree(h, tmp);
    dfree(h, queues);
    ctr = nullptr;
    tmp = queues = nullptr;
  };
  if (rc != NF_OK) {
    free_scratch();
    return rc;
  }
  hipEvent_t* ev = h->ev;
  hipEventRecord(ev[0], s);
  fill_u32((uint32_t*)ctr, 0u, 16, s);
  fill_u32(queues, 0u, 2 * QWORDS, s);
  PairArgs a{};
  SegCrossArgs xa{};
  // ---- bin: centre cell list
  if (cells) {
    rc = build_contact_cells(h);
  } else {
    double edge = 0.0;
    int fine = 1;
    footprint_cells(h, n, &edge, &fine);
    c.fp_edge = edge;
    c.fp_fine = fine;
    xa.g = seg_grid(p.box, edge, (double)amax, gap, ew);
    const double rpad = (double)xa.g.r + 2.0 * (double)CELL_MARGIN;
    rc = build_cell_list(h, edge, (double)hlmax * REACH_PAD + rpad, false, SEG_ROW);
    if (rc == NF_OK) c.cell_list = 2;
    xa.fine = fine;
    xa.rpad = (float)rpad;
    const float rr = (float)((double)hlmax * REACH_PAD + rpad) * 1.0001f;
    xa.reach2 = rr * rr;
  }
  if (rc != NF_OK) {
    free_scratch();
    return rc;
  }
  if (cells) {
    a.desc = nullptr;
    a.st = nullptr;
    a.cell_inst = nullptr;
    a.ncells_total = c.ncells;
    a.stencil = c.stencil;
    a.st_off = 0;
    a.st_n = c.st_n;
    a.prP = c.cP;
    a.prD = c.cD;
    a.prX = c.cid;
    a.prStart = c.cstart;
    a.prCount = c.ccount;
    a.nbP = c.cP;
    a.nbD = c.cD;
    a.nbX = c.cid;
    a.nbStart = c.cstart;
    a.nbCount = c.ccount;
    a.nbHl = c.chl;
    a.qctr = queues;
    for (int k = 0; k < 3; ++k) {
      a.box[k] = p.box[k];
      a.boxf[k] = (float)p.box[k];
      a.ncell[k] = c.ncell[k];
    }
    a.gap64 = gap;
    a.gap32 = (float)gap;
    a.ew = (float)ew;
    a.hl_max = hlmax;
    a.R_max = (float)Rmax * 1.000001f;
    a.ctr = ctr;
    a.diag = h->opt_diag;
  } else {
    // the claim grid is the cell list's grid
    for (int k = 0; k < 3; ++k) {
      xa.g.nc[k] = c.ncell[k];
      xa.g.h[k] = (float)(p.box[k] / (double)c.ncell[k]);
      xa.g.inv_h[k] = (float)((double)c.ncell[k] / p.box[k]);
      xa.g.inv_h64[k] = (double)c.ncell[k] / p.box[k];
      xa.x.box[k] = p.box[k];
      xa.x.boxf[k] = (float)p.box[k];
    }
    xa.g.ncells = c.ncells;
    xa.cP = c.cP;
    xa.cD = c.cD;
    xa.cid = c.cid;
    xa.cstart = c.cstart;
    xa.ccount = c.ccount;
    xa.chl = c.chl;
    xa.stencil = c.stencil;
    xa.st_n = c.st_n;
    xa.x.gap64 = gap;
    xa.x.gap32 = (float)gap;
    xa.x.ew = (float)ew;
    xa.x.diag = h->opt_diag;
    xa.ctr = ctr;
    xa.next = queues;  // eight counters, QSTRIDE words apart
    xa.overflow = (uint32_t*)(ctr + 6);
    // each rod's crossing of its neighbour cells, once per rod
    rc = dalloc(h, &c.cmask, std::max<uint64_t>(n, 1));
    if (rc != NF_OK) {
      free_scratch();
      return rc;
    }
    launch_seg_marks(c.cP, c.cD, n, c.ncell, c.cscale, xa.g.h, xa.g.r, 4.0f * xa.x.ew, c.cmask, s);
    xa.cmask = c.cmask;
  }
  hipEventRecord(ev[1], s);
  // one pair pass appends every contact to an unsorted list.  Append
  // capacity: the previous call's count, else one contact per two rods; a
  // pass that overflows is repeated with the exact size
  uint64_t cap = std::max<uint64_t>(h->contact_hint + h->contact_hint / 8 + 1024, (uint64_t)n / 2 + 4096);
  if (cap > 0xFFFFFFF0ull) cap = 0xFFFFFFF0ull;
  uint32_t* appI = nullptr;
  uint32_t* appJ = nullptr;
  float* appG = nullptr;
  double* appG64 = nullptr;
  int32_t* appImg = nullptr;
  auto free_app = [&]() {
    dfree(h, appI);
    dfree(h, appJ);
    dfree(h, appG);
    dfree(h, appG64);
    dfree(h, appImg);
    appI = appJ = nullptr;
    appG = nullptr;
    appG64 = nullptr;
    appImg = nullptr;
  };
  auto free_all = [&]() {
    free_app();
    free_scratch();
  };
  uint32_t total = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    rc = NF_OK;
    if (rc == NF_OK) rc = dalloc(h, &appI, cap);
    if (rc == NF_OK) rc = dalloc(h, &appJ, cap);
    if (rc == NF_OK) rc = dalloc(h, &appG, cap);
    if (rc == NF_OK) rc = dalloc(h, &appG64, cap);
    if (rc == NF_OK) rc = dalloc(h, &appImg, cap);
    if (rc != NF_OK) {
      free_all();
      return rc;
    }
    if (attempt > 0) {
      fill_u32((uint32_t*)ctr, 0u, 16, s);
      fill_u32(queues, 0u, 2 * QWORDS, s);
    }
    hipEventRecord(ev[1], s);
    if (cells) {
      a.app_n = (uint32_t*)(ctr + 4);
      a.app_cap = (uint32_t)cap;
      a.outI = appI;
      a.outJ = appJ;
      a.outGap = appG;
      a.outGap64 = appG64;
      a.outImg = appImg;
      launch_pair(PM_CNT, prec, a, s);
    } else {
      xa.app_n = (uint32_t*)(ctr + 4);
      xa.app_cap = (uint32_t)cap;
      xa.outI = appI;
      xa.outJ = appJ;
      xa.outGap = appG;
      xa.outGap64 = appG64;
      xa.outImg = appImg;
      launch_seg_cross(prec, xa, s);
    }
    hipEventRecord(ev[2], s);
    trace("contacts: %u rods, %s pass launched", n, cells ? "cell-pair" : "footprint");
    uint32_t over = 0;
    HIPCHK(h, hipMemcpyAsync(&total, (uint32_t*)(ctr + 4), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipMemcpyAsync(&over, (uint32_t*)(ctr + 6), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(h, sync_stream(s, "contact pair pass"));
    HIPCHK(h, hipGetLastError());
    if (over) {
      trace("contacts: a home cell has more survivors than fit; cell-pair engine instead");
      free_all();
      return 1;
    }
    if ((uint64_t)total <= cap) break;
    trace("contacts: %u contacts exceed the append capacity %llu, repeating the pass", total,
          (unsigned long long)cap);
    free_app();
    cap = (uint64_t)total;
    if (attempt == 1) {
      free_all();
      return fail(h, NF_E_STATE, "contact count changed between two identical passes");
    }
  }
  h->contact_hint = total;
  c.n_contacts = total;
  c.cap_contacts = total;
  rc = NF_OK;
  if (rc == NF_OK) rc = dalloc(h, &c.J, total);
  if (rc == NF_OK) rc = dalloc(h, &c.gap, total);
  if (rc == NF_OK) rc = dalloc(h, &c.gap64, total);
  if (rc == NF_OK) rc = dalloc(h, &c.img, total);
  if (rc == NF_OK) rc = dalloc(h, &c.area, total);
  if (rc != NF_OK) {
    free_all();
    return rc;
  }
  // CSR over I: counts, scan, scatter, canonical row order
  hipEventRecord(ev[3], s);
  fill_u32(c.cnt, 0u, (uint64_t)n + 1, s);
  fill_u32(c.fill, 0u, n, s);
  launch_csr_count(appI, total, c.cnt, s);
  scan_u32(c.cnt, c.rowptr, n, tmp, s);
  launch_csr_scatter(appI, appJ, appG, appG64, appImg, total, c.rowptr, c.fill, c.J, c.gap, c.gap64,
                     c.img, s);
  hipEventRecord(ev[4], s);
  launch_row_sort(c.rowptr, n, c.J, c.gap, c.gap64, c.img, s);
  launch_contact_area(P, D, n, c.rowptr, c.J, c.img, p.box, gap, c.area, s);
  hipEventRecord(ev[5], s);
  HIPCHK(h, hipMemcpyAsync(h->cb_stats, ctr, sizeof(unsigned long long) * 6, hipMemcpyDeviceToHost, s));
  HIPCHK(h, sync_stream(s, "contact list"));
  trace("contacts: %u recorded and sorted", total);
  HIPCHK(h, hipGetLastError());
  float ms = 0.0f;
  const int from[4] = {0, 1, 3, 4}, to[4] = {1, 2, 4, 5};
  for (int k = 0; k < 4; ++k) {
    (void)hipEventElapsedTime(&ms, ev[from[k]], ev[to[k]]);
    h->cb_ms[k] = ms;
  }
  free_all();
  c.valid = true;
  c.engine = cells ? NF_ENGINE_CELLS : NF_ENGINE_FOOTPRINT;
  if (cells) {
    c.fp_edge = 0.0;
    c.fp_fine = 0;
  }
  if (n_out) *n_out = total;
  return NF_OK;
}

int contacts_run(nf_handle* h, double gap, uint64_t* n_out, int prec, unsigned long long*) {
  const bool cells = h->opt_contact_engine == NF_ENGINE_CELLS;
  int rc = contacts_pass(h, gap, n_out, prec, cells);
  if (rc == 1) rc = contacts_pass(h, gap, n_out, prec, true);
  return rc;
}

}  // namespace nf

extern "C" int nf_contacts(nf_handle* h, double gap, uint64_t* n_contacts) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  h->err.clear();
  if (!(gap >= 0.0)) return fail(h, NF_E_INVALID, "gap must be >= 0");
  const int prec = h->opt_contact_prec >= 0 ? h->opt_contact_prec : h->params[h->post_inst].precision;
  return contacts_run(h, gap, n_contacts, prec, nullptr);
}

extern "C" int nf_get_contacts(nf_handle* h, uint32_t* row_ptr, uint32_t* col, float* gap,
                               int32_t* image, uint64_t cap) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  ContactBuffers& c = h->cb;
  if (!c.valid) return fail(h, NF_E_STATE, "no contacts computed");
  if ((col || gap || image) && cap < c.n_contacts) return fail(h, NF_E_CAPACITY, "contact buffer too small");
  if (row_ptr)
    HIPCHK(h, hipMemcpy(row_ptr, c.rowptr, sizeof(uint32_t) * ((size_t)c.n_rods + 1), hipMemcpyDeviceToHost));
  if (c.n_contacts) {
    if (col) HIPCHK(h, hipMemcpy(col, c.J, sizeof(uint32_t) * c.n_contacts, hipMemcpyDeviceToHost));
    if (gap) HIPCHK(h, hipMemcpy(gap, c.gap, sizeof(float) * c.n_contacts, hipMemcpyDeviceToHost));
    if (image) HIPCHK(h, hipMemcpy(image, c.img, sizeof(int32_t) * c.n_contacts, hipMemcpyDeviceToHost));
  }
  return NF_OK;
}

extern "C" int nf_get_contact_gap64(nf_handle* h, double* gap64, uint64_t cap) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  ContactBuffers& c = h->cb;
  if (!c.valid) return fail(h, NF_E_STATE, "no contacts computed");
  if (cap < c.n_contacts) return fail(h, NF_E_CAPACITY, "buffer too small");
  if (c.n_contacts && gap64)
    HIPCHK(h, hipMemcpy(gap64, c.gap64, sizeof(double) * c.n_contacts, hipMemcpyDeviceToHost));
  return NF_OK;
}

extern "C" int nf_get_contact_area(nf_handle* h, double* area, uint64_t cap) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  ContactBuffers& c = h->cb;
  if (!c.valid) return fail(h, NF_E_STATE, "no contacts computed");
  if (cap < c.n_contacts) return fail(h, NF_E_CAPACITY, "buffer too small");
  if (c.n_contacts && area)
    HIPCHK(h, hipMemcpy(area, c.area, sizeof(double) * c.n_contacts, hipMemcpyDeviceToHost));
  return NF_OK;
}

extern "C" int nf_contacts_stats(nf_handle* h, uint64_t* pair_tests, uint64_t* rechecks,
                                 double* max_dd, float* kernel_ms, uint64_t* tile_bytes) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  if (!h->cb.valid) return fail(h, NF_E_STATE, "no contacts computed");
  if (pair_tests) *pair_tests = h->cb_stats[0];
  if (rechecks) *rechecks = h->cb_stats[1];
  if (max_dd) {
    const unsigned long long b = h->cb_stats[2];
    double v;
    std::memcpy(&v, &b, sizeof(v));
    *max_dd = v;
  }
  if (kernel_ms)
    for (int k = 0; k < 4; ++k) kernel_ms[k] = h->cb_ms[k];
  if (tile_bytes) *tile_bytes = h->cb_stats[3];
  return NF_OK;
}

extern "C" int nf_contacts_engine(nf_handle* h, int* engine, double* home_edge, int* fine,
                                  uint64_t* split_units) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  if (!h->cb.valid) return fail(h, NF_E_STATE, "no contacts computed");
  if (engine) *engine = h->cb.engine;
  if (home_edge) *home_edge = h->cb.fp_edge;
  if (fine) *fine = h->cb.fp_fine;
  if (split_units) *split_units = h->cb.engine == NF_ENGINE_FOOTPRINT ? h->cb_stats[5] : 0;
  return NF_OK;
}

// ===========================================================================
// histograms
// ===========================================================================
namespace nf {

int hist_prepare(nf_handle* h, const nf_hist_spec* spec, bool want_mutual, bool want_2d,
                 double* edges, HistPlan* plan) {
  if (!h || !spec) return fail(h, NF_E_INVALID, "null argument");
  const int nb = spec->bins, npb = spec->phi_bins, ntb = spec->theta_bins;
  if (nb < 1 || nb > HIST_BINS_MAX) return fail(h, NF_E_INVALID, "bins must be 1..256");
  if (spec->centre_axis < 0 || spec->centre_axis > 2) return fail(h, NF_E_INVALID, "centre_axis must be 0..2");
  const bool two_d = npb > 0 && want_2d;
  if (two_d && (npb % 2 != 0 || ntb < 1 || npb * ntb > HIST_2D_MAX))
    return fail(h, NF_E_INVALID, "phi-theta needs even phi_bins and phi*theta <= 4096");
  if (want_mutual && !h->cb.valid) return fail(h, NF_E_STATE, "mutual-angle histogram needs nf_contacts first");
  const double PI = 3.14159265358979323846;
  // edges: polar [nb+1], mutual [nb+1], then theta [ntb+1] and phi [half+1]
  size_t k0 = 0;
  for (int k = 0; k <= nb; ++k) edges[k0 + k] = std::cos((double)k * PI / (double)nb);
  k0 += (size_t)nb + 1;
  for (int k = 0; k <= nb; ++k) edges[k0 + k] = std::cos((double)k * (0.5 * PI) / (double)nb);
  k0 += (size_t)nb + 1;
  if (two_d) {
    const int half = npb / 2;
    for (int k = 0; k <= ntb; ++k) edges[k0 + k] = std::cos((double)k * PI / (double)ntb);
    k0 += (size_t)ntb + 1;
    for (int k = 0; k <= half; ++k) edges[k0 + k] = std::cos((double)k * PI / (double)half);
    k0 += (size_t)half + 1;
  }
  double mu[3] = {spec->pol_axis[0], spec->pol_axis[1], spec->pol_axis[2]};
  const double mn = std::sqrt((mu[0] * mu[0] + mu[1] * mu[1]) + mu[2] * mu[2]);
  if (!(mn > 0.0)) return fail(h, NF_E_INVALID, "pol_axis must be non-zero");
  for (int k = 0; k < 3; ++k) plan->mu[k] = mu[k] / mn;
  plan->nb = nb;
  plan->npb = two_d ? npb : 0;
  plan->ntb = two_d ? ntb : 0;
  plan->axis = spec->centre_axis;
  plan->mutual = want_mutual;
  plan->n_edges = k0;
  plan->n_out = 3 * (size_t)nb + (two_d ? (size_t)npb * ntb : 0);
  return NF_OK;
}

int hist_run(nf_handle* h, const HistPlan& pl, const double* edges, double* dE,
             unsigned long long* dOut, int64_t* out_centre, int64_t* out_polar, int64_t* out_mutual,
             int64_t* out_phitheta) {
  hipStream_t s = h->stream;
  const int nb = pl.nb;
  hipError_t e = hipMemcpyAsync(dE, edges, sizeof(double) * pl.n_edges, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    fill_u32((uint32_t*)dOut, 0u, 2 * pl.n_out, s);
    e = hipGetLastError();
  }
  const InstDesc& d = h->desc[h->post_inst];
  const uint32_t n = h->n_rods[h->post_inst];
  if (e == hipSuccess && n) {
    launch_hist(h->gb.bP + d.rod_off, h->gb.bD + d.rod_off, n, pl.mutual ? h->cb.rowptr : nullptr,
                h->cb.J, dE, dE + (nb + 1), pl.npb ? dE + 2 * (nb + 1) : nullptr, nb, pl.npb, pl.ntb,
                pl.axis, h->params[h->post_inst].box[pl.axis], pl.mu, dOut, s);
    e = hipGetLastError();
  }
  // int64 and u64 share the layout; counts never reach 2^63
  const int64_t* src = (const int64_t*)dOut;
  struct {
    int64_t* dst;
    size_t off, len;
  } cp[4] = {{out_centre, 0, (size_t)nb},
             {out_polar, (size_t)nb, (size_t)nb},
             {pl.mutual ? out_mutual : nullptr, 2 * (size_t)nb, (size_t)nb},
             {pl.npb ? out_phitheta : nullptr, 3 * (size_t)nb, (size_t)pl.npb * pl.ntb}};
  for (auto& c : cp)
    if (e == hipSuccess && c.dst)
      e = hipMemcpyAsync(c.dst, src + c.off, sizeof(int64_t) * c.len, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(h, NF_E_HIP, std::string("histograms: ") + hipGetErrorString(e));
  return NF_OK;
}

}  // namespace nf

extern "C" int nf_histograms(nf_handle* h, const nf_hist_spec* spec, int64_t* out_centre,
                             int64_t* out_polar, int64_t* out_mutual, int64_t* out_phitheta) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  double edges[HIST_EDGES_MAX];
  HistPlan pl;
  int rc = hist_prepare(h, spec, out_mutual != nullptr, out_phitheta != nullptr, edges, &pl);
  if (rc != NF_OK) return rc;
  double* dE = nullptr;
  unsigned long long* dOut = nullptr;
  rc = dalloc(h, &dE, pl.n_edges);
  if (rc == NF_OK) rc = dalloc(h, &dOut, pl.n_out);
  if (rc == NF_OK)
    rc = hist_run(h, pl, edges, dE, dOut, out_centre, out_polar, out_mutual, out_phitheta);
  dfree(h, dE);
  dfree(h, dOut);
  return rc;
}

// Centre-centre distance statistics of instance 0 (statKernel_CENT via
// collectStats, src/statistics_collection.cu:866-1033).
extern "C" int nf_centre_distance(nf_handle* h, int32_t bins, double width, int64_t* counts,
                                  uint64_t* n_pairs, uint64_t* sum_d_q24, uint64_t* sum_d2_q16) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  if (bins < 1 || bins > CENT_BINS_MAX || !(width > 0.0))
    return fail(h, NF_E_INVALID, "need 1 <= bins <= 1024 and width > 0");
  const nf_params& p = h->params[h->post_inst];
  const uint32_t n = h->n_rods[h->post_inst];
  const double rmax = bins * width;
  const double bmin = std::min(p.box[0], std::min(p.box[1], p.box[2]));
  if (rmax > 8.0 * bmin) return fail(h, NF_E_INVALID, "bins * width exceeds 8 box edges");
  hipStream_t s = h->stream;
  const float4* P = h->gb.bP + h->desc[h->post_inst].rod_off;
  const float4* D = h->gb.bD + h->desc[h->post_inst].rod_off;
  // grid: edge about rmax (half stencil of +-1 cells), finer when crowded
  const double edge = default_cell_edge(rmax, (double)n / (p.box[0] * p.box[1] * p.box[2]), 1.0);
  int32_t nc[3];
  double scale[3];
  uint64_t ncells = 1;
  for (int a = 0; a < 3; ++a) {
    double k = std::floor(p.box[a] / edge);
    if (k < 1.0) k = 1.0;
    if (k > 1024.0) k = 1024.0;
    nc[a] = (int32_t)k;
    scale[a] = (double)nc[a] / p.box[a];
    ncells *= (uint64_t)nc[a];
  }
  if (ncells > (1ull << 26)) return fail(h, NF_E_INVALID, "statistics grid too large");
  std::vector<int4> st = build_stencil(nc, p.box, rmax, true);
  std::vector<double> e2(bins + 1);
  for (int k = 0; k <= bins; ++k) e2[k] = ((double)k * width) * ((double)k * width);
  float4* cP = nullptr;
  float4* cD = nullptr;
  uint32_t *cid = nullptr, *cstart = nullptr, *ccount = nullptr;
  float* chl = nullptr;
  int4* dst = nullptr;
  double* de2 = nullptr;
  unsigned long long* out = nullptr;
  int rc = NF_OK;
  const uint64_t n1 = std::max<uint64_t>(n, 1);
  if (rc == NF_OK) rc = dalloc(h, &cP, n1);
  if (rc == NF_OK) rc = dalloc(h, &cD, n1);
  if (rc == NF_OK) rc = dalloc(h, &cid, n1);
  if (rc == NF_OK) rc = dalloc(h, &cstart, ncells + 1);
  if (rc == NF_OK) rc = dalloc(h, &ccount, ncells);
  if (rc == NF_OK) rc = dalloc(h, &chl, ncells);
  if (rc == NF_OK) rc = dalloc(h, &dst, st.size());
  if (rc == NF_OK) rc = dalloc(h, &de2, (size_t)bins + 1);
  if (rc == NF_OK) rc = dalloc(h, &out, (size_t)bins + 3);
  BinScratch bs;
  auto release = [&]() {
    void* w[] = {cP, cD, cid, cstart, ccount, chl, dst, de2, out};
    for (void* q : w) dfree(h, q);
    free_bin_scratch(h, &bs);
  };
  if (rc != NF_OK) {
    release();
    return rc;
  }
  hipError_t e = hipMemcpyAsync(dst, st.data(), sizeof(int4) * st.size(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(de2, e2.data(), sizeof(double) * e2.size(), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    release();
    return fail(h, NF_E_HIP, std::string("statistics upload: ") + hipGetErrorString(e));
  }
  fill_u32((uint32_t*)out, 0u, 2 * ((uint64_t)bins + 3), s);
  rc = bin_rods(h, P, D, n, nc, scale, cP, cD, cid, cstart, ccount, chl, &bs);
  if (rc != NF_OK) {
    (void)sync_stream(s, "centre distance");
    release();
    return rc;
  }
  const double Lmax = std::max(p.box[0], std::max(p.box[1], p.box[2]));
  const float ew = (float)(5.9604644775390625e-8 * (Lmax + rmax) * 1.01);
  const float rf = (float)rmax + 8.0f * ew;
  launch_cent(cP, cid, cstart, ccount, dst, (uint32_t)st.size(), nc, p.box, de2, bins,
              rf * rf * 1.00001f, out, s);
  std::vector<unsigned long long> hout((size_t)bins + 3);
  e = hipMemcpyAsync(hout.data(), out, sizeof(unsigned long long) * hout.size(), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = sync_stream(s, "centre distance");
  if (e == hipSuccess) e = hipGetLastError();
  release();
  if (e != hipSuccess) return fail(h, NF_E_HIP, std::string("centre distance: ") + hipGetErrorString(e));
  if (counts)
    for (int k = 0; k < bins; ++k) counts[k] = (int64_t)hout[k];
  if (n_pairs) *n_pairs = hout[bins];
  if (sum_d_q24) *sum_d_q24 = hout[bins + 1];
  if (sum_d2_q16) *sum_d2_q16 = hout[bins + 2];
  return NF_OK;
}

// Van der Waals energy over the contacts of the last nf_contacts call
// (collect_vdw, src/statistics_collection.cu:53-154).
extern "C" int nf_vdw(nf_handle* h, double A, double B, double* energy, uint64_t* n_pairs) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  ContactBuffers& c = h->cb;
  if (!c.valid) return fail(h, NF_E_STATE, "no contacts computed");
  if (!(c.gap_cut > 0.0)) return fail(h, NF_E_INVALID, "the contact gap must be > 0");
  hipStream_t s = h->stream;
  const uint32_t n = c.n_rods;
  const float4* P = h->gb.bP + h->desc[h->post_inst].rod_off;
  const float4* D = h->gb.bD + h->desc[h->post_inst].rod_off;
  double *rowsum = nullptr, *bsum = nullptr, *tot = nullptr;
  unsigned long long* pairs = nullptr;
  int rc = NF_OK;
  const uint64_t nb = ((uint64_t)n + 1023) / 1024;
  if (rc == NF_OK) rc = dalloc(h, &rowsum, std::max<uint64_t>(n, 1));
  if (rc == NF_OK) rc = dalloc(h, &bsum, std::max<uint64_t>(nb, 1));
  if (rc == NF_OK) rc = dalloc(h, &tot, 1);
  if (rc == NF_OK) rc = dalloc(h, &pairs, 1);
  auto release = [&]() {
    dfree(h, rowsum);
    dfree(h, bsum);
    dfree(h, tot);
    dfree(h, pairs);
  };
  if (rc != NF_OK) {
    release();
    return rc;
  }
  fill_u32((uint32_t*)pairs, 0u, 2, s);
  launch_vdw(P, D, n, c.rowptr, c.J, c.img, h->params[h->post_inst].box, c.gap_cut, A, B, rowsum, bsum, tot,
             pairs, s);
  double e_h = 0.0;
  unsigned long long p_h = 0;
  hipError_t e = hipMemcpyAsync(&e_h, tot, sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&p_h, pairs, sizeof(p_h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = sync_stream(s, "van der Waals");
  if (e == hipSuccess) e = hipGetLastError();
  release();
  if (e != hipSuccess) return fail(h, NF_E_HIP, std::string("van der Waals: ") + hipGetErrorString(e));
  if (energy) *energy = e_h;
  if (n_pairs) *n_pairs = p_h;
  return NF_OK;
}
