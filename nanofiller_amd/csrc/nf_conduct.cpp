// nf_conduct.cpp - host side of the DC conductance solve on the network of
// the last nf_network call (see nf_solve.hip): reduced system, conjugate
// gradients with a convergence check every CG_CHECK iterations, currents and
// rod potentials.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "nf_host.h"

using namespace nf;

namespace {
constexpr uint32_t CG_CHECK = 16;  // iterations queued between host convergence checks
}

extern "C" int nf_conductance(nf_handle* h, double tol, uint32_t max_iter, nf_conductance_info* info,
                              double* potentials) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  ContactBuffers& c = h->cb;
  if (!c.net_valid) return fail(h, NF_E_STATE, "no network built (call nf_network first)");
  if (!(tol > 0.0) || max_iter == 0) return fail(h, NF_E_INVALID, "tol must be > 0 and max_iter > 0");
  const uint32_t n = c.n_rods;
  hipStream_t s = h->stream;
  uint32_t *flag = nullptr, *act = nullptr, *rl = nullptr, *idx = nullptr, *off = nullptr,
           *tmp = nullptr, *crow = nullptr, *ccol = nullptr;
  double *cval = nullptr, *dinv = nullptr, *bsrc = nullptr, *bsnk = nullptr, *x = nullptr,
         *r = nullptr, *z = nullptr, *pv = nullptr, *q = nullptr, *part = nullptr, *vdev = nullptr;
  CgScal* scal = nullptr;
  uint32_t *acto = nullptr, *rlo = nullptr, *idxo = nullptr, *offo = nullptr;
  auto release = [&]() {
    void* ptrs[] = {flag, act, rl,   idx, off, tmp,  crow, ccol, cval, dinv, bsrc, bsnk,
                    x,    r,   z,    pv,  q,   part, vdev, scal, acto, rlo,  idxo, offo};
    for (void* p : ptrs) dfree(h, p);
  };
  const uint64_t n1 = (uint64_t)n + 1;
  int rc = NF_OK;
#define ALLOC(ptr, cnt) \
  if (rc == NF_OK) rc = dalloc(h, &(ptr), (size_t)std::max<uint64_t>((cnt), 1));
  ALLOC(flag, n);
  ALLOC(act, n1);
  ALLOC(rl, n1);
  ALLOC(idx, n1);
  ALLOC(off, n1);
  ALLOC(tmp, scan_tmp_words(n1));
  ALLOC(part, 2 * CG_MAX_BLOCKS);
  ALLOC(scal, 1);
  if (rc != NF_OK) {
    release();
    return rc;
  }
  // reduced system: rods of the components touching both electrodes
  fill_u32(flag, 0u, std::max<uint64_t>(n, 1), s);
  fill_u32(act, 0u, n1, s);
  fill_u32(rl, 0u, n1, s);
  if (n > 0) {
    launch_cs_flag(c.nrow, c.ncol, c.labels, n, flag, s);
    launch_cs_active(c.nrow, c.ncol, c.labels, flag, n, act, rl, s);
  }
  // unknowns numbered in the cell list's (spatial) order when the packing's
  // cell list is at hand, else in rod order
  const bool spatial = c.cid != nullptr && c.cell_list != 0 && n > 0;
  if (spatial) {
    if (rc == NF_OK) rc = dalloc(h, &acto, n1);
    if (rc == NF_OK) rc = dalloc(h, &rlo, n1);
    if (rc == NF_OK) rc = dalloc(h, &idxo, n1);
    if (rc == NF_OK) rc = dalloc(h, &offo, n1);
    if (rc != NF_OK) {
      release();
      return rc;
    }
    fill_u32(acto + n, 0u, 1, s);
    fill_u32(rlo + n, 0u, 1, s);
    launch_cs_order(n, c.cid, act, rl, acto, rlo, nullptr, nullptr, nullptr, nullptr, 0, s);
    scan_u32(acto, idxo, n1, tmp, s);
    scan_u32(rlo, offo, n1, tmp, s);
    launch_cs_order(n, c.cid, nullptr, nullptr, nullptr, nullptr, idxo, offo, idx, off, 1, s);
  } else {
    scan_u32(act, idx, n1, tmp, s);
    scan_u32(rl, off, n1, tmp, s);
  }
  uint32_t mn[2] = {0, 0};
  if (hipError_t e_ = hipMemcpyAsync(&mn[0], idx + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s); e_ != hipSuccess) {
      release();
      return fail(h, NF_E_HIP, std::string("conductance copy: ") + hipGetErrorString(e_));
    }
  if (hipError_t e_ = hipMemcpyAsync(&mn[1], off + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s); e_ != hipSuccess) {
      release();
      return fail(h, NF_E_HIP, std::string("conductance copy: ") + hipGetErrorString(e_));
    }
  if (hipError_t e = sync_stream(s, "conductance set-up"); e != hipSuccess) {
    release();
    return fail(h, NF_E_HIP, std::string("conductance set-up: ") + hipGetErrorString(e));
  }
  {
    void* q[] = {acto, rlo, idxo, offo};
    for (void* p : q) dfree(h, p);
    acto = rlo = idxo = offo = nullptr;
  }
  const uint32_t m = mn[0], nnz = mn[1];
  nf_conductance_info out{};
  out.unknowns = m;
  out.converged = 1;
  float solve_ms = 0.0f;
  if (m > 0) {
    ALLOC(crow, (uint64_t)m + 1);
    ALLOC(ccol, nnz);
    ALLOC(cval, nnz);
    ALLOC(dinv, m);
    ALLOC(bsrc, m);
    ALLOC(bsnk, m);
    ALLOC(x, m);
    ALLOC(r, m);
    ALLOC(z, m);
    ALLOC(pv, m);
    ALLOC(q, m);
    if (rc != NF_OK) {
      release();
      return rc;
    }
    launch_cs_fill(c.nrow, c.ncol, c.nval, n, act, idx, off, crow, ccol, cval, dinv, bsrc, bsnk, s);
    fill_u32(crow + m, nnz, 1, s);
    CgArgs a{m, crow, ccol, cval, dinv, bsrc, bsnk, x, r, z, pv, q, part, scal};
    hipEventRecord(h->ev[0], s);
    launch_cg_init(a, s);
    uint32_t it = 0;
    CgScal hs{};
    bool conv = false;
    for (;;) {
      if (hipError_t e_ = hipMemcpyAsync(&hs, scal, sizeof(CgScal), hipMemcpyDeviceToHost, s); e_ != hipSuccess) {
      release();
      return fail(h, NF_E_HIP, std::string("conductance copy: ") + hipGetErrorString(e_));
    }
      if (hipError_t e = sync_stream(s, "conductance iterations"); e != hipSuccess) {
        release();
        return fail(h, NF_E_HIP, std::string("conductance iterations: ") + hipGetErrorString(e));
      }
      const double res = hs.bb > 0.0 ? std::sqrt(std::max(hs.rr, 0.0) / hs.bb) : 0.0;
      out.residual = res;
      if (res <= tol) {
        conv = true;
        break;
      }
      if (it >= max_iter || !std::isfinite(res)) break;
      const uint32_t k = std::min(CG_CHECK, max_iter - it);
      for (uint32_t j = 0; j < k; ++j) launch_cg_iteration(a, s);
      it += k;
    }
    launch_cg_current(a, s);
    hipEventRecord(h->ev[1], s);
    if (hipError_t e_ = hipMemcpyAsync(&hs, scal, sizeof(CgScal), hipMemcpyDeviceToHost, s); e_ != hipSuccess) {
      release();
      return fail(h, NF_E_HIP, std::string("conductance copy: ") + hipGetErrorString(e_));
    }
    if (hipError_t e = sync_stream(s, "conductance currents"); e != hipSuccess) {
      release();
      return fail(h, NF_E_HIP, std::string("conductance currents: ") + hipGetErrorString(e));
    }
    hipEventElapsedTime(&solve_ms, h->ev[0], h->ev[1]);
    out.iterations = it;
    out.converged = conv ? 1 : 0;
    out.conductance = hs.isrc;
    out.current_sink = hs.isnk;
  }
  out.solve_ms = solve_ms;
  if (potentials && n > 0) {
    ALLOC(vdev, n);
    if (rc != NF_OK) {
      release();
      return rc;
    }
    launch_cs_potential(n, c.labels, flag, act, idx, x, vdev, s);
    if (hipError_t e_ = hipMemcpyAsync(potentials, vdev, sizeof(double) * n, hipMemcpyDeviceToHost, s); e_ != hipSuccess) {
      release();
      return fail(h, NF_E_HIP, std::string("conductance copy: ") + hipGetErrorString(e_));
    }
    if (hipError_t e = sync_stream(s, "conductance potentials"); e != hipSuccess) {
      release();
      return fail(h, NF_E_HIP, std::string("conductance potentials: ") + hipGetErrorString(e));
    }
  }
#undef ALLOC
  release();
  if (info) *info = out;
  return NF_OK;
}
