// nf_api.cpp - C ABI of the library: parameter normalisation, handle
// lifecycle, batched generation (HIP-graph rounds), contacts, histograms,
// copies out.  Host orchestration only; all per-rod and per-pair work runs
// in the HIP kernels of nf_gen.hip, nf_pair.hip and nf_post.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <thread>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nf_internal.h"
#include "nf_host.h"

using namespace nf;

std::string g_last_error;

namespace nf {

int fail(nf_handle* h, int code, const std::string& msg) {
  if (h)
    h->err = msg;
  else
    g_last_error = msg;
  return code;
}

// Size classes: 256 B granules below 64 KiB, then four classes per power of
// two, so a reused block wastes at most a quarter of its size.
static size_t size_class(size_t bytes) {
  if (bytes <= 65536) return (bytes + 255) & ~(size_t)255;
  size_t p = 65536;
  while (p < bytes) p <<= 1;
  const size_t q = p >> 3;  // p/2 + k*p/8, k = 1..4
  size_t c = p >> 1;
  while (c < bytes) c += q;
  return c;
}

constexpr size_t POOL_CAP = (size_t)16 << 30;  // freed bytes kept per handle

int dalloc_bytes(nf_handle* h, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  const size_t cls = size_class(bytes);
  auto it = h->pool.find(cls);
  void* q = nullptr;
  if (it != h->pool.end()) {
    q = it->second;
    h->pool.erase(it);
    h->pool_bytes -= cls;
  } else {
    hipError_t e = hipMalloc(&q, cls);
    if (e != hipSuccess && !h->pool.empty()) {
      // give the cached blocks back and try once more
      for (auto& kv : h->pool) (void)hipFree(kv.second);
      h->pool.clear();
      h->pool_bytes = 0;
      e = hipMalloc(&q, cls);
    }
    if (e != hipSuccess)
      return fail(h, NF_E_NOMEM,
                  "hipMalloc of " + std::to_string(cls) + " bytes failed: " + hipGetErrorString(e));
  }
  h->allocs[q] = cls;
  *p = q;
  return NF_OK;
}

static bool trace_on() {
  static int on = -1;
  if (on < 0) {
    const char* e = std::getenv("NF_TRACE");
    on = (e && *e && *e != '0') ? 1 : 0;
  }
  return on == 1;
}

void trace(const char* fmt, ...) {
  if (!trace_on()) return;
  va_list ap;
  va_start(ap, fmt);
  std::fprintf(stderr, "[nf] ");
  std::vfprintf(stderr, fmt, ap);
  std::fprintf(stderr, "\n");
  std::fflush(stderr);
  va_end(ap);
}

hipError_t sync_stream(hipStream_t s, const char* what) {
  static double limit = -1.0;
  if (limit < 0.0) {
    const char* e = std::getenv("NF_SYNC_TIMEOUT_S");
    limit = e ? std::atof(e) : 120.0;
    if (!(limit > 0.0)) limit = 120.0;
  }
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    const hipError_t e = hipStreamQuery(s);
    if (e != hipErrorNotReady) return e;
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (el > limit) {
      std::fprintf(stderr, "[nf] %s: stream not done after %.0f s\n", what, el);
      std::fflush(stderr);
      return hipErrorLaunchTimeOut;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(el < 0.01 ? 2 : 200));
  }
}

// Returns a block to the handle's pool.  The handle's stream is drained
// first (usually already idle), so no queued kernel can still use a block
// that a later call receives.
void dfree(nf_handle* h, void* p) {
  if (!p) return;
  auto it = h->allocs.find(p);
  if (it == h->allocs.end()) {
    (void)hipFree(p);
    return;
  }
  const size_t cls = it->second;
  h->allocs.erase(it);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->pool_bytes + cls > POOL_CAP) {
    (void)hipFree(p);
    return;
  }
  h->pool.emplace(cls, p);
  h->pool_bytes += cls;
}

}  // namespace nf

// ===========================================================================
// parameters
// ===========================================================================
extern "C" void nf_params_default(nf_params* p) {
  std::memset(p, 0, sizeof(*p));
  // simPrms::simPrms (src/simPrms.cpp:22-117)
  p->vol_frac = 0.05 / 100.0;
  p->len_law = NF_LEN_CONST;
  p->len = 4.0;
  p->len_sigma = 0.0;
  p->len_min = 0.0;
  p->len_max = 0.0;
  p->radius = 0.010;
  p->sep = 0.0;
  p->soft_frac = 1.0;
  p->ort_law = NF_ORT_ISO;
  p->kappa = 0.0;
  p->axis[0] = 0.0;
  p->axis[1] = 0.0;
  p->axis[2] = 1.0;
  p->den_law = NF_DEN_UNIFORM;
  p->agg_sigma = 1.0;
  p->agg_peak = 1.0;
  p->contact_gap = 0.015;
  p->max_rounds = 1000;  // cudaMain.cu:113
  p->stall_rounds = 8;
  p->min_batch = 64;
  p->max_batch = 0;
  p->cell_size = 0.0;
  p->precision = NF_PREC_FP32_RECHECK;
  p->elec_axis = 0;
  p->tunnel_lambda = 1e-4;
  p->tunnel_g0 = 1.0;
  p->seed = 1;
  p->stream = 0;
}

double nf::default_cell_edge(double reach_typ, double density, double factor) {
  double e = factor * reach_typ;
  if (density > 0.0 && density * e * e * e > 256.0) {
    e = std::cbrt(32.0 / density);
    if (e < 0.25 * reach_typ) e = 0.25 * reach_typ;
    if (e > reach_typ) e = reach_typ;
  }
  return e;
}

static double excluded_volume(double L, double a) {
  const double D = 2.0 * a;
  const double PI = 3.14159265358979323846;
  return 0.5 * PI * D * L * L + 2.0 * PI * D * D * L + (4.0 * PI / 3.0) * D * D * D;
}

extern "C" int nf_params_normalize(nf_params* p) {
  for (int d = 0; d < 3; ++d) {
    if (!(p->box[d] > 0.0)) return fail(nullptr, NF_E_INVALID, "box edges must be > 0");
    p->box[d] = (double)(float)p->box[d];
  }
  if (!(p->radius > 0.0)) return fail(nullptr, NF_E_INVALID, "radius must be > 0");
  if (!(p->len > 0.0)) return fail(nullptr, NF_E_INVALID, "length must be > 0");
  if (p->len_law == NF_LEN_LOGNORMAL) {
    if (!(p->len_min > 0.0) || !(p->len_max >= p->len_min) || !(p->len_sigma >= 0.0))
      return fail(nullptr, NF_E_INVALID, "log-normal length needs 0 < len_min <= len_max, sigma >= 0");
  } else if (p->len_law != NF_LEN_CONST) {
    return fail(nullptr, NF_E_INVALID, "unknown length law");
  }
  if (p->ort_law == NF_ORT_VMF) {
    if (!(p->kappa > 0.0)) return fail(nullptr, NF_E_INVALID, "vMF kappa must be > 0");
    const double n = std::sqrt((p->axis[0] * p->axis[0] + p->axis[1] * p->axis[1]) + p->axis[2] * p->axis[2]);
    if (!(n > 0.0)) return fail(nullptr, NF_E_INVALID, "vMF axis must be non-zero");
    for (int d = 0; d < 3; ++d) p->axis[d] = p->axis[d] / n;
  } else if (p->ort_law != NF_ORT_ISO) {
    return fail(nullptr, NF_E_INVALID, "unknown orientation law");
  }
  if (p->den_law == NF_DEN_AGGLOM) {
    if (p->n_agg < 1 || !(p->agg_sigma > 0.0) || !(p->agg_peak >= 1.0))
      return fail(nullptr, NF_E_INVALID, "agglomerates need n_agg >= 1, sigma > 0, peak >= 1");
  } else if (p->den_law == NF_DEN_MESH) {
    if (!p->den_mesh) return fail(nullptr, NF_E_INVALID, "the mesh density law needs den_mesh");
    uint64_t nv = 1;
    for (int d = 0; d < 3; ++d) {
      if (p->den_ext[d] < 1 || p->den_ext[d] > 4096)
        return fail(nullptr, NF_E_INVALID, "density map extents must be 1..4096");
      nv *= (uint64_t)p->den_ext[d];
    }
    double sum = 0.0;
    for (uint64_t i = 0; i < nv; ++i) {
      const float v = p->den_mesh[i];
      if (!(v >= 0.0f && v <= 1.0f))
        return fail(nullptr, NF_E_INVALID, "density map values must lie in [0, 1]");
      sum += v;
    }
    if (!(sum > 0.0)) return fail(nullptr, NF_E_INVALID, "density map is zero everywhere");
  } else if (p->den_law != NF_DEN_UNIFORM) {
    return fail(nullptr, NF_E_INVALID, "unknown density law");
  }
  if (!(p->soft_frac > 0.0) || p->sep < 0.0) return fail(nullptr, NF_E_INVALID, "bad hard core");
  {
    // periodic images per axis stay bounded: rod reach at most 16 box edges
    const double lmax = p->len_law == NF_LEN_LOGNORMAL ? p->len_max : p->len;
    const double reach = lmax + 2.0 * p->radius * p->soft_frac + p->sep + p->contact_gap;
    const double bmin = std::min(p->box[0], std::min(p->box[1], p->box[2]));
    if (reach > 16.0 * bmin) return fail(nullptr, NF_E_INVALID, "rods longer than 16 box edges");
  }
  const double V = p->box[0] * p->box[1] * p->box[2];
  const double PI = 3.14159265358979323846;
  if (p->n_rods == 0) {
    // simPrms::set_numCNT (src/simPrms.cpp:987-1002): N = ceil(vf V / (pi a^2 l))
    const double vcnt = PI * p->radius * p->radius * p->len;
    const double n = std::ceil(p->vol_frac * V / vcnt);
    if (!(n >= 1.0) || n > 4.0e9) return fail(nullptr, NF_E_INVALID, "rod count out of range");
    p->n_rods = (uint32_t)n;
  }
  if (p->min_batch == 0) p->min_batch = 64;
  if (p->max_batch == 0) {
    const double vex = excluded_volume(p->len, p->radius);
    // three excluded volumes' worth of candidates per round: on average a
    // candidate's excluded volume holds three others of its round.  Fewer,
    // larger rounds cost less than the extra same-round rejections
    // (measured on MI355X: 500k-rod packing 4.77 -> 4.17 ms, 5 -> 3 rounds;
    // 8M-rod packing 103 -> 91 ms, 19 -> 9 rounds; the 256-packing batch
    // 281 -> 290 ms)
    double m = 3.0 * V / vex;
    const double cap = 4.0 * (double)p->n_rods + 1024.0;
    if (m > cap) m = cap;
    if (m > (double)(1u << 23)) m = (double)(1u << 23);
    if (m < 256.0) m = 256.0;
    p->max_batch = (uint32_t)m;
  }
  if (p->min_batch > p->max_batch) p->min_batch = p->max_batch;
  if (p->max_rounds == 0) p->max_rounds = 1000;
  if (p->stall_rounds == 0) p->stall_rounds = 8;
  if (p->precision != NF_PREC_FP64 && p->precision != NF_PREC_FP32_RECHECK)
    return fail(nullptr, NF_E_INVALID, "unknown precision");
  if (p->elec_axis < 0 || p->elec_axis > 2) return fail(nullptr, NF_E_INVALID, "elec_axis must be 0..2");
  return NF_OK;
}

// ===========================================================================
// stencils
// ===========================================================================
namespace nf {
std::vector<int4> build_stencil(const int32_t nc[3], const double box[3], double reach, bool half, int bx) {
  std::vector<std::pair<float, int4>> v;
  const double rm = reach + (double)CELL_MARGIN;
  if (bx < 1 || (half && bx != 1)) bx = 1;
  int k[3];
  double E[3];
  for (int d = 0; d < 3; ++d) {
    E[d] = box[d] / nc[d];
    k[d] = (int)std::floor(rm / E[d]) + 1;
  }
  for (int dz = -k[2]; dz <= k[2]; ++dz)
    for (int dy = -k[1]; dy <= k[1]; ++dy)
      for (int dx = -k[0]; dx <= k[0] + bx - 1; ++dx) {
        if (half) {
          const bool pos = dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx >= 0)));
          if (!pos) continue;
        }
        // whole cells between the offset cell and the block, per axis
        const int gap[3] = {dx >= bx ? dx - bx : (dx < 0 ? -dx - 1 : 0), std::abs(dy) - 1, std::abs(dz) - 1};
        double m2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          const int a = gap[d];
          if (a > 0) m2 += (a * E[d]) * (a * E[d]);
        }
        if (m2 <= rm * rm) {
          // store a slightly low minimum distance so float rounding never prunes
          const float md2 = (float)(m2 * (1.0 - 1e-6));
          int4 e;
          e.x = dx;
          e.y = dy;
          e.z = dz;
          e.w = 0;
          std::memcpy(&e.w, &md2, sizeof(float));
          v.push_back({md2, e});
        }
      }
  std::stable_sort(v.begin(), v.end(), [](const std::pair<float, int4>& a, const std::pair<float, int4>& b) {
    return a.first < b.first;
  });
  std::vector<int4> out;
  out.reserve(v.size());
  for (auto& e : v) out.push_back(e.second);
  return out;
}
}  // namespace nf

// ===========================================================================
// lifecycle
// ===========================================================================
static int plan_instance(const nf_params& p, InstDesc& d, std::vector<double>& agg,
                         std::vector<int4>& stencil) {
  std::memset(&d, 0, sizeof(d));
  SampleDesc& s = d.smp;
  for (int a = 0; a < 3; ++a) s.box[a] = p.box[a];
  s.len = p.len;
  s.len_sigma = p.len_sigma;
  s.len_min = p.len_min;
  s.len_max = p.len_max;
  s.radius = p.radius;
  s.sep = p.sep;
  s.soft_frac = p.soft_frac;
  s.kappa = p.kappa;
  for (int a = 0; a < 3; ++a) s.axis[a] = p.axis[a];
  if (p.ort_law == NF_ORT_VMF) vmf_basis(s.axis, s.e1, s.e2);
  s.agg_sigma = p.agg_sigma;
  s.agg_peak = p.agg_peak;
  s.seed = p.seed;
  s.stream = p.stream;
  s.len_law = p.len_law;
  s.ort_law = p.ort_law;
  s.den_law = p.den_law;
  s.n_agg = p.den_law == NF_DEN_AGGLOM ? p.n_agg : 0;
  s.agg = nullptr;
  s.mesh = nullptr;  // device copy, set at upload
  for (int a = 0; a < 3; ++a) s.mext[a] = p.den_law == NF_DEN_MESH ? p.den_ext[a] : 0;
  agg.resize(3 * (size_t)s.n_agg);
  for (int k = 0; k < s.n_agg; ++k) agg_centre(p.seed, p.stream, p.box, (uint32_t)k, &agg[3 * k]);

  const float lmaxf = (float)(p.len_law == NF_LEN_LOGNORMAL ? p.len_max : p.len);
  const float af = (float)p.radius;
  d.hl_max = 0.5f * lmaxf;
  const double Rmax = p.soft_frac * ((double)af + (double)af) + p.sep;
  d.R_max = (float)Rmax * 1.000001f;
  const double reach = ((double)d.hl_max + (double)d.hl_max) * REACH_PAD + Rmax;
  // cell edge: requested, or about one typical reach; when that would put
  // more than 256 rods in an average cell (long rods, small boxes) the edge
  // drops towards ~32 rods per cell, but not below a quarter of the reach
  // (a negative cell_size scales that default edge)
  double target = p.cell_size;
  const double V = p.box[0] * p.box[1] * p.box[2];
  if (!(target > 0.0)) {
    // 0.85 of a typical reach: the reject-vs-accepted pass stages the rods
    // of every cell within reach of the home cell, so slightly smaller cells
    // stage fewer far rods (measured optimum on the 500k-rod workload)
    const double e = default_cell_edge(p.len * REACH_PAD + Rmax, (double)p.n_rods / V, 0.85);
    target = target < 0.0 ? -target * e : e;
  }
  uint64_t ncells = 1;
  for (int a = 0; a < 3; ++a) {
    double n = std::floor(p.box[a] / target);
    if (n < 1.0) n = 1.0;
    if (n > 1024.0) n = 1024.0;
    d.ncell[a] = (int32_t)n;
    d.boxf[a] = (float)p.box[a];
    d.cell_scale[a] = (double)d.ncell[a] / p.box[a];
    d.edge[a] = (float)(p.box[a] / d.ncell[a]);
    ncells *= (uint64_t)d.ncell[a];
  }
  if (ncells > (1ull << 26)) return NF_E_INVALID;
  d.ncells = (uint32_t)ncells;
  // full stencil, then the half stencil of the same-round pass
  stencil = build_stencil(d.ncell, p.box, reach, false);
  d.st_n = (uint32_t)stencil.size();
  const std::vector<int4> half = build_stencil(d.ncell, p.box, reach, true);
  d.hst_n = (uint32_t)half.size();
  stencil.insert(stencil.end(), half.begin(), half.end());
  // per-cell capacity for accepted rods
  const double cellV = V / (double)ncells;
  double peak = p.den_law == NF_DEN_AGGLOM ? p.agg_peak : 1.0;
  if (p.den_law == NF_DEN_MESH) {
    // densest point relative to the mean of the map
    const uint64_t nv = (uint64_t)p.den_ext[0] * p.den_ext[1] * p.den_ext[2];
    double sum = 0.0, mx = 0.0;
    for (uint64_t i = 0; i < nv; ++i) {
      sum += p.den_mesh[i];
      mx = std::max(mx, (double)p.den_mesh[i]);
    }
    peak = mx * (double)nv / sum;
  }
  const double lam = (double)p.n_rods * cellV / V * peak;
  double cap = std::ceil(lam * 1.25 + 8.0 * std::sqrt(lam) + 16.0);
  if (cap > (double)p.n_rods + 16.0) cap = (double)p.n_rods + 16.0;
  d.cap = (uint32_t)cap;
  d.n_target = p.n_rods;
  d.min_batch = p.min_batch;
  d.max_batch = p.max_batch;
  d.max_rounds = p.max_rounds;
  d.stall_rounds = p.stall_rounds;
  d.sf64 = p.soft_frac;
  d.sep64 = p.sep;
  d.sf32 = (float)p.soft_frac;
  d.sep32 = (float)p.sep;
  const double Lmax = std::max(p.box[0], std::max(p.box[1], p.box[2]));
  d.ew = (float)(5.9604644775390625e-8 * (Lmax + reach) * 1.01);
  return NF_OK;
}

static void release_gen(nf_handle* h) {
  GenBuffers& b = h->gb;
  void* ptrs[] = {b.desc, b.st, b.ri, b.slot_base, b.stencil, b.agg, b.mesh, b.cell_inst, b.oP, b.oD,
                  b.oid, b.ostart, b.ocount, b.ohl, b.bP, b.bD, b.sP, b.sD, b.scell, b.srk,
                  b.surv, b.srank, b.pP, b.pD, b.pslot, b.palive, b.pcount, b.pstart, b.phl,
                  b.order, b.aP, b.aD, b.aslot, b.aidx, b.acount, b.astart, b.queues, b.scan_tmp};
  for (void* p : ptrs) dfree(h, p);
  b = GenBuffers();
}

void nf::release_contacts(nf_handle* h) {
  ContactBuffers& c = h->cb;
  void* ptrs[] = {c.cP, c.cD, c.cid, c.cstart, c.ccount, c.chl, c.cmask, c.stencil, c.cnt,
                  c.rowptr, c.fill, c.J, c.gap, c.gap64, c.img, c.area, c.labels, c.nrow, c.ncol, c.nval,
                  c.nd_pflag, c.nd_edge_a, c.nd_edge_b, c.nd_esrc, c.nd_esnk, c.nd_src_node,
                  c.nd_snk_node, c.nd_egs, c.nd_egk, c.nd_node_rod, c.nd_node_t, c.nd_seg_a,
                  c.nd_seg_b, c.nd_seg_L, c.nd_seg_A, c.nd_ce_a, c.nd_ce_b, c.nd_ce_k, c.nd_cel_node,
                  c.nd_cel_gap, c.nd_cel_side};
  for (void* p : ptrs) dfree(h, p);
  c = ContactBuffers();
}

extern "C" int nf_create_batch(const nf_params* ps, int n, nf_handle** out) {
  if (!out) return fail(nullptr, NF_E_INVALID, "null output handle");
  *out = nullptr;
  if (!ps || n < 1 || n > 65536) return fail(nullptr, NF_E_INVALID, "need 1..65536 instances");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(nullptr, NF_E_NODEVICE, "no HIP device visible");
  nf_handle* h = new nf_handle();
  h->params.assign(ps, ps + n);
  for (int i = 0; i < n; ++i) {
    int rc = nf_params_normalize(&h->params[i]);
    if (rc != NF_OK) {
      std::string m = "instance " + std::to_string(i) + ": " + g_last_error;
      delete h;
      return fail(nullptr, rc, m);
    }
  }
  (void)hipGetDevice(&h->device);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(nullptr, NF_E_HIP, "hipStreamCreate failed");
  }
  for (auto& e : h->ev) (void)hipEventCreate(&e);

  // plan all instances
  h->desc.resize(n);
  std::vector<std::vector<double>> aggs(n);
  std::vector<std::vector<int4>> sts(n);
  GenBuffers& b = h->gb;
  b.n_inst = n;
  uint64_t cell_base = 0, old_off = 0, rod_off = 0, T_cap = 0, agg_total = 0, st_total = 0,
           mesh_total = 0;
  for (int i = 0; i < n; ++i) {
    InstDesc& d = h->desc[i];
    if (plan_instance(h->params[i], d, aggs[i], sts[i]) != NF_OK) {
      delete h;
      return fail(nullptr, NF_E_INVALID, "instance " + std::to_string(i) + ": cell grid too large");
    }
    d.cell_base = (uint32_t)cell_base;
    d.old_off = old_off;
    d.rod_off = (uint32_t)rod_off;
    d.st_off = (uint32_t)st_total;
    d.hst_off = (uint32_t)st_total + d.st_n;
    cell_base += d.ncells;
    old_off += (uint64_t)d.ncells * d.cap;
    rod_off += d.n_target;
    T_cap += d.max_batch;
    agg_total += aggs[i].size();
    if (h->params[i].den_law == NF_DEN_MESH)
      mesh_total += (uint64_t)d.smp.mext[0] * d.smp.mext[1] * d.smp.mext[2];
    st_total += sts[i].size();
    if (cell_base > 0xFFFFFFF0ull || old_off > 0xFFFFFFF0ull || rod_off > 0xFFFFFFF0ull ||
        T_cap > 0xFFFFFFF0ull) {
      delete h;
      return fail(nullptr, NF_E_INVALID, "batch exceeds 32-bit index space");
    }
  }
  b.total_cells = cell_base;
  b.total_old = old_off;
  b.total_rods = rod_off;
  b.T_cap = T_cap;
  int rc = NF_OK;
#define ALLOC(p, n_)                      \
  if (rc == NF_OK) rc = dalloc(h, &(p), (n_));
  ALLOC(b.desc, n);
  ALLOC(b.st, n);
  ALLOC(b.ri, 1);
  ALLOC(b.slot_base, n + 1);
  ALLOC(b.stencil, st_total);
  ALLOC(b.agg, agg_total);
  ALLOC(b.mesh, mesh_total);
  ALLOC(b.cell_inst, b.total_cells);
  ALLOC(b.oP, b.total_old);
  ALLOC(b.oD, b.total_old);
  ALLOC(b.oid, b.total_old);
  ALLOC(b.ostart, b.total_cells);
  ALLOC(b.ocount, b.total_cells);
  ALLOC(b.ohl, b.total_cells);
  ALLOC(b.bP, b.total_rods);
  ALLOC(b.bD, b.total_rods);
  ALLOC(b.sP, b.T_cap);
  ALLOC(b.sD, b.T_cap);
  ALLOC(b.scell, b.T_cap);
  ALLOC(b.srk, b.T_cap);
  ALLOC(b.surv, b.T_cap + 1);
  ALLOC(b.srank, b.T_cap + 1);
  ALLOC(b.pP, b.T_cap);
  ALLOC(b.pD, b.T_cap);
  ALLOC(b.pslot, b.T_cap);
  ALLOC(b.palive, b.T_cap);
  ALLOC(b.pcount, b.total_cells);
  ALLOC(b.pstart, b.total_cells + 1);
  ALLOC(b.order, b.total_cells);
  ALLOC(b.aP, b.T_cap);
  ALLOC(b.aD, b.T_cap);
  ALLOC(b.aslot, b.T_cap);
  ALLOC(b.aidx, b.T_cap + 1);
  ALLOC(b.acount, b.total_cells);
  ALLOC(b.astart, b.total_cells + 1);
  ALLOC(b.phl, b.total_cells);
  ALLOC(b.queues, Q_N * QWORDS);
  ALLOC(b.scan_tmp, std::max(scan_tmp_words(b.total_cells), scan_tmp_words(b.T_cap)));
#undef ALLOC
  if (rc != NF_OK) {
    std::string m = h->err;
    nf_destroy(h);
    return fail(nullptr, rc, m);
  }
  // upload descriptors, agglomerates, stencils, per-cell tables
  std::vector<double> agg_all;
  std::vector<float> mesh_all;
  std::vector<int4> st_all;
  std::vector<uint32_t> cinst(b.total_cells), ostart(b.total_cells);
  for (int i = 0; i < n; ++i) {
    InstDesc& d = h->desc[i];
    d.smp.agg = aggs[i].empty() ? nullptr : b.agg + agg_all.size();
    agg_all.insert(agg_all.end(), aggs[i].begin(), aggs[i].end());
    if (h->params[i].den_law == NF_DEN_MESH) {
      const uint64_t nv = (uint64_t)d.smp.mext[0] * d.smp.mext[1] * d.smp.mext[2];
      d.smp.mesh = b.mesh + mesh_all.size();
      mesh_all.insert(mesh_all.end(), h->params[i].den_mesh, h->params[i].den_mesh + nv);
      h->params[i].den_mesh = nullptr;  // the caller's buffer: not kept past nf_create
    }
    st_all.insert(st_all.end(), sts[i].begin(), sts[i].end());
    for (uint32_t c = 0; c < d.ncells; ++c) {
      cinst[d.cell_base + c] = (uint32_t)i;
      ostart[d.cell_base + c] = (uint32_t)(d.old_off + (uint64_t)c * d.cap);
    }
  }
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMemcpy(b.desc, h->desc.data(), sizeof(InstDesc) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess && !agg_all.empty())
    e = hipMemcpy(b.agg, agg_all.data(), sizeof(double) * agg_all.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !mesh_all.empty())
    e = hipMemcpy(b.mesh, mesh_all.data(), sizeof(float) * mesh_all.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !st_all.empty())
    e = hipMemcpy(b.stencil, st_all.data(), sizeof(int4) * st_all.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b.cell_inst, cinst.data(), sizeof(uint32_t) * b.total_cells, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b.ostart, ostart.data(), sizeof(uint32_t) * b.total_cells, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    nf_destroy(h);
    return fail(nullptr, NF_E_HIP, std::string("upload failed: ") + hipGetErrorString(e));
  }
  h->n_rods.assign(n, 0);
  h->planned_cap.resize(n);
  for (int i = 0; i < n; ++i) h->planned_cap[i] = h->desc[i].cap;
  *out = h;
  return NF_OK;
}

extern "C" int nf_create(const nf_params* p, nf_handle** out) { return nf_create_batch(p, 1, out); }

extern "C" void nf_destroy(nf_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->graph) (void)hipGraphExecDestroy(h->graph);
  release_contacts(h);
  release_gen(h);
  for (auto& kv : h->allocs) (void)hipFree(kv.first);
  for (auto& kv : h->pool) (void)hipFree(kv.second);
  for (auto& e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

extern "C" const char* nf_last_error(const nf_handle* h) {
  return h ? h->err.c_str() : g_last_error.c_str();
}

extern "C" int nf_num_instances(const nf_handle* h) { return h ? (int)h->params.size() : 0; }

extern "C" const char* nf_version(void) { return "nanofiller-mi355x 0.1 (gfx950)"; }

static int scale_slots(nf_handle* h, double scale);

// A captured round graph holds the launch shapes and buffers of its capture.
static void drop_graph(nf_handle* h) {
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
}

extern "C" int nf_set_option(nf_handle* h, int key, int64_t value) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  switch (key) {
    case NF_OPT_GRAPH: h->opt_graph = value ? 1 : 0; break;
    case NF_OPT_PROFILE: h->opt_profile = value ? 1 : 0; break;
    case NF_OPT_DIAG: h->opt_diag = value ? 1 : 0; break;
    case NF_OPT_CONTACT_PREC:
      if (value != -1 && value != NF_PREC_FP64 && value != NF_PREC_FP32_RECHECK)
        return fail(h, NF_E_INVALID, "bad contact precision");
      h->opt_contact_prec = (int)value;
      break;
    case NF_OPT_CONTACT_ENGINE:
      if (value != NF_ENGINE_FOOTPRINT && value != NF_ENGINE_CELLS)
        return fail(h, NF_E_INVALID, "bad contact engine");
      h->opt_contact_engine = (int)value;
      break;
    case NF_OPT_FOOTPRINT_FINE:
      if (value < 0 || value > 1000000) return fail(h, NF_E_INVALID, "footprint fine cell factor out of range");
      h->opt_footprint_fine = (int)value;
      break;
    case NF_OPT_WAVES_OLD:
    case NF_OPT_WAVES_NEW:
      if (value < 0 || value > 16) return fail(h, NF_E_INVALID, "waves per CU out of range (0..16)");
      (key == NF_OPT_WAVES_OLD ? h->gb.tiles.waves_old : h->gb.tiles.waves_new) = (uint32_t)value;
      drop_graph(h);
      break;
    case NF_OPT_PART_SHIFT:
      if (value < -1 || value > 3) return fail(h, NF_E_INVALID, "part shift out of range (-1..3)");
      h->gb.tiles.part_shift = (int32_t)value;
      drop_graph(h);
      break;
    case NF_OPT_SPLIT_MIN:
      if (value < 0 || value > 4096) return fail(h, NF_E_INVALID, "split minimum out of range (0..4096)");
      h->gb.tiles.split_min = (uint32_t)value;
      drop_graph(h);
      break;
    case NF_OPT_TEST_ALL_PAIRS:
      h->gb.tiles.test_all = value ? 1u : 0u;
      drop_graph(h);
      break;
    case NF_OPT_SLOT_SCALE:
      if (value < 1 || value > 1000000) return fail(h, NF_E_INVALID, "slot scale out of range (1..1000000)");
      return scale_slots(h, (double)value * 1e-3);
    case NF_OPT_POST_INSTANCE:
      if (value < 0 || value >= (int64_t)h->params.size()) return fail(h, NF_E_INVALID, "no such instance");
      if ((int)value != h->post_inst) {
        release_contacts(h);  // contacts, network and cell lists belong to the previous instance
        h->post_inst = (int)value;
      }
      break;
    case NF_OPT_FOOTPRINT_OCC:
      if (value < 0 || value > 200) return fail(h, NF_E_INVALID, "footprint occupancy out of range (0..200)");
      h->opt_footprint_occ = (int)value;
      break;
    default: return fail(h, NF_E_INVALID, "unknown option");
  }
  return NF_OK;
}

// ===========================================================================
// generation
// ===========================================================================
static int reset_generation(nf_handle* h) {
  GenBuffers& b = h->gb;
  HIPCHK(h, hipMemsetAsync(b.st, 0, sizeof(InstState) * b.n_inst, h->stream));
  HIPCHK(h, hipMemsetAsync(b.ri, 0, sizeof(RoundInfo), h->stream));
  fill_u32(b.ocount, 0u, b.total_cells, h->stream);
  fill_u32((uint32_t*)b.ohl, 0u, b.total_cells, h->stream);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, sync_stream(h->stream, "generation reset"));
  return NF_OK;
}

static int build_graph(nf_handle* h, int prec, int diag) {
  if (h->graph && h->graph_prec == prec && h->graph_diag == diag) return NF_OK;
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  launch_round(h->gb, prec, diag, h->stream, nullptr);
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (e != hipSuccess) return fail(h, NF_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(h, NF_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  h->graph_prec = prec;
  h->graph_diag = diag;
  return NF_OK;
}

// Milestone dumps (VOLUME_PERCENT_LIST, src/cudaMain.cu:136-142): after a
// round, instance 0's volume fraction from its accepted count; when it has
// reached the next listed value, its packing so far (the accepted rods sit
// at the front of its rod buffer, in acceptance order) is saved as a state
// file named by the fraction reached, and every value reached is passed.
// The reference compares the fraction in percent with the list converted
// to fractions (a hundredfold early trigger) and reads past the list's end
// after its last value; here both are percent and the list ends the dumps.
static int milestone_check(nf_handle* h, size_t* next) {
  if (*next >= h->milestones.size()) return NF_OK;
  InstState st0{};
  HIPCHK(h, hipMemcpy(&st0, h->gb.st, sizeof(InstState), hipMemcpyDeviceToHost));
  const nf_params& p = h->params[0];
  const double PI = 3.14159265358979323846;
  const double vf = (double)st0.n_acc * PI * p.radius * p.radius * p.len / (p.box[0] * p.box[1] * p.box[2]) * 100.0;
  if (vf < h->milestones[*next]) return NF_OK;
  char name[64];
  std::snprintf(name, sizeof(name), "state_%4.2f.bin", vf);
  const std::string path = h->milestone_dir.empty() ? std::string(name) : h->milestone_dir + "/" + name;
  const uint32_t keep = h->n_rods[0];
  h->n_rods[0] = st0.n_acc;
  const int rc = nf_save_state(h, path.c_str());
  h->n_rods[0] = keep;
  if (rc != NF_OK) return rc;
  trace("generate: milestone %s (%u rods)", path.c_str(), st0.n_acc);
  while (*next < h->milestones.size() && h->milestones[*next] <= vf) ++*next;
  return NF_OK;
}

extern "C" int nf_set_milestones(nf_handle* h, const double* vf_percent, int32_t n, const char* dir) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  if (n < 0 || (n > 0 && !vf_percent)) return fail(h, NF_E_INVALID, "bad milestone list");
  if (n > 0 && h->params.size() != 1) return fail(h, NF_E_INVALID, "milestones need a single-instance handle");
  std::vector<double> v(vf_percent, vf_percent + n);
  for (size_t i = 0; i < v.size(); ++i)
    if (!(v[i] > 0.0) || (i > 0 && !(v[i] > v[i - 1])))
      return fail(h, NF_E_INVALID, "milestones must be positive and ascending");
  h->milestones = v;
  h->milestone_dir = dir ? dir : "";
  return NF_OK;
}

// Per-cell slots of the accepted rods (the "old" set): instance i gets
// caps[i] slots per cell.  Reallocates the slot arrays, rewrites each
// instance's offset and every cell's first slot, and drops a captured round
// graph (its kernel arguments name the old arrays).  The packing does not
// depend on the capacity, only whether it suffices.
static int relayout_old(nf_handle* h, const std::vector<uint32_t>& caps) {
  GenBuffers& b = h->gb;
  const int n = (int)h->desc.size();
  uint64_t off = 0;
  for (int i = 0; i < n; ++i) {
    InstDesc& d = h->desc[i];
    d.cap = caps[i];
    d.old_off = off;
    off += (uint64_t)d.ncells * d.cap;
  }
  if (off > 0xFFFFFFF0ull) return fail(h, NF_E_INVALID, "cell slots exceed the 32-bit index space");
  dfree(h, b.oP);
  dfree(h, b.oD);
  dfree(h, b.oid);
  b.oP = b.oD = nullptr;
  b.oid = nullptr;
  int rc = dalloc(h, &b.oP, std::max<uint64_t>(off, 1));
  if (rc == NF_OK) rc = dalloc(h, &b.oD, std::max<uint64_t>(off, 1));
  if (rc == NF_OK) rc = dalloc(h, &b.oid, std::max<uint64_t>(off, 1));
  if (rc != NF_OK) return rc;
  b.total_old = off;
  std::vector<uint32_t> ostart(b.total_cells);
  for (int i = 0; i < n; ++i) {
    const InstDesc& d = h->desc[i];
    for (uint32_t c = 0; c < d.ncells; ++c) ostart[d.cell_base + c] = (uint32_t)(d.old_off + (uint64_t)c * d.cap);
  }
  HIPCHK(h, hipMemcpy(b.desc, h->desc.data(), sizeof(InstDesc) * n, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(b.ostart, ostart.data(), sizeof(uint32_t) * b.total_cells, hipMemcpyHostToDevice));
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return NF_OK;
}

// The planned slots per cell scaled by scale (at least one, at most what
// cannot overflow: the instance's target + 16).
static int scale_slots(nf_handle* h, double scale) {
  std::vector<uint32_t> caps(h->desc.size());
  for (size_t i = 0; i < caps.size(); ++i) {
    const double c = std::ceil((double)h->planned_cap[i] * scale);
    caps[i] = (uint32_t)std::max(1.0, std::min(c, (double)h->desc[i].n_target + 16.0));
  }
  return relayout_old(h, caps);
}

static int generate_once(nf_handle* h, nf_gen_stats* stats, std::vector<int>* overflowed);

// Generation with slot growth: an instance whose cell slots overflowed is
// generated again from the start with twice the slots per cell (the packing
// is the same; only its storage was too small), up to the capacity that
// cannot overflow.
extern "C" int nf_generate(nf_handle* h, nf_gen_stats* stats) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  uint32_t growths = 0;
  for (;;) {
    std::vector<int> over;
    const int rc = generate_once(h, stats, &over);
    if (rc != NF_E_OVERFLOW || over.empty()) {
      if (stats)
        for (int i = 0; i < h->gb.n_inst; ++i) stats[i].slot_growths = growths;
      return rc;
    }
    std::vector<uint32_t> caps(h->desc.size());
    bool grew = false;
    for (size_t i = 0; i < caps.size(); ++i) caps[i] = h->desc[i].cap;
    for (int i : over) {
      const uint32_t lim = h->desc[i].n_target + 16u;
      const uint32_t c = (uint32_t)std::min<uint64_t>(2ull * caps[i], lim);
      grew = grew || c > caps[i];
      caps[i] = c;
    }
    if (!grew) return rc;
    const int r2 = relayout_old(h, caps);
    if (r2 != NF_OK) return r2;
    ++growths;
    trace("generate: cell slots overflowed, regenerating with %u slots per cell (instance %d)", caps[over[0]],
          over[0]);
  }
}

static int generate_once(nf_handle* h, nf_gen_stats* stats, std::vector<int>* overflowed) {
  h->err.clear();
  const int prec = h->params[0].precision;
  for (auto& p : h->params)
    if (p.precision != prec) return fail(h, NF_E_INVALID, "all instances of a batch share one precision");
  int rc = reset_generation(h);
  if (rc) return rc;
  const bool use_graph = h->opt_graph && !h->opt_profile;
  trace("generate: %d instance(s), precision %d, graph %d, profile %d", h->gb.n_inst, prec,
        (int)use_graph, h->opt_profile);
  if (use_graph) {
    rc = build_graph(h, prec, h->opt_diag);
    if (rc) return rc;
    trace("generate: graph built");
  }
  double kern[8] = {0};
  RoundInfo ri{};
  RoundInfo* ri_host = nullptr;
  HIPCHK(h, hipHostMalloc((void**)&ri_host, sizeof(RoundInfo), hipHostMallocDefault));
  uint32_t max_rounds = 0;
  for (auto& p : h->params) max_rounds = std::max(max_rounds, p.max_rounds);
  const auto t0 = std::chrono::steady_clock::now();
  uint32_t launched = 0;
  int batch = 1;
  std::vector<std::array<hipEvent_t, 8>> prof_ev;
  size_t prof_used = 0;
  size_t next_ms = 0;  // next milestone (volume fraction) to dump
  hipError_t e = hipSuccess;
  for (;;) {
    for (int k = 0; k < batch && e == hipSuccess; ++k) {
      if (use_graph) {
        e = hipGraphLaunch(h->graph, h->stream);
      } else {
        hipEvent_t* ev = nullptr;
        if (h->opt_profile) {
          // one event set per round, read back after the final sync
          if (prof_used == prof_ev.size()) {
            prof_ev.emplace_back();
            for (auto& x : prof_ev.back()) (void)hipEventCreate(&x);
          }
          ev = prof_ev[prof_used++].data();
        }
        launch_round(h->gb, prec, h->opt_diag, h->stream, ev);
        e = hipGetLastError();
      }
      ++launched;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(ri_host, h->gb.ri, sizeof(RoundInfo), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = sync_stream(h->stream, "generation round");
    if (e != hipSuccess) break;
    ri = *ri_host;
    trace("generate: %u rounds launched, controller rounds %u, running %u, last T %u", launched,
          ri.rounds, ri.running, ri.T);
    if (!h->milestones.empty()) {
      const int mrc = milestone_check(h, &next_ms);
      if (mrc != NF_OK) {
        (void)hipHostFree(ri_host);
        return mrc;
      }
    }
    if (ri.running == 0 || launched > max_rounds + 2) break;
    // next batch: the rounds the controller expects to need at the last
    // round's acceptance (at most 8 between host checks); one at a time
    // while a milestone is pending
    batch = (int)std::min<uint32_t>(std::max<uint32_t>(ri.est, 1u), 8u);
    if (next_ms < h->milestones.size()) batch = 1;
  }
  const auto t1 = std::chrono::steady_clock::now();
  (void)hipHostFree(ri_host);
  if (e == hipSuccess && h->opt_profile) {
    const int cls[7] = {NF_K_CONTROL, NF_K_GEN, NF_K_BIN, NF_K_REJECT_OLD, NF_K_REJECT_NEW,
                        NF_K_COMMIT, NF_K_CONTROL};
    for (size_t r = 0; r < prof_used; ++r)
      for (int c = 0; c < 7; ++c) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, prof_ev[r][c], prof_ev[r][c + 1]) == hipSuccess) kern[cls[c]] += ms;
      }
  }
  for (auto& a : prof_ev)
    for (auto& x : a) (void)hipEventDestroy(x);
  if (e != hipSuccess) return fail(h, NF_E_HIP, std::string("generation: ") + hipGetErrorString(e));
  const double wall = std::chrono::duration<double, std::milli>(t1 - t0).count();
  std::vector<InstState> st(h->gb.n_inst);
  HIPCHK(h, hipMemcpy(st.data(), h->gb.st, sizeof(InstState) * st.size(), hipMemcpyDeviceToHost));
  int status_rc = NF_OK;
  for (int i = 0; i < h->gb.n_inst; ++i) {
    h->n_rods[i] = st[i].n_acc;
    if (st[i].status == NF_STATUS_OVERFLOW) {
      status_rc = NF_E_OVERFLOW;
      overflowed->push_back(i);
    }
    if (stats) {
      nf_gen_stats& s = stats[i];
      std::memset(&s, 0, sizeof(s));
      s.n_target = h->desc[i].n_target;
      s.n_accepted = st[i].n_acc;
      s.rounds = st[i].round;
      s.status = (int32_t)st[i].status;
      s.candidates = st[i].candidates;
      s.proposals = st[i].proposals;
      s.pair_tests = st[i].pair_tests;
      s.rechecks = st[i].rechecks;
      s.rejected_old = st[i].rej_old;
      s.rejected_new = st[i].rej_new;
      s.wall_ms = wall;
      for (int k = 0; k < 8; ++k) s.kern_ms[k] = kern[k];
      s.max_abs_d32_d64 = st[i].max_dd;
      s.tile_bytes = st[i].tile_bytes;
      s.tile_bytes_new = st[i].tile_bytes_new;
      s.rounds_launched = launched;
    }
  }
  h->generated = true;
  h->cb.valid = false;
  h->cb.net_valid = false;
  h->cb.nd_valid = false;
  if (status_rc != NF_OK) return fail(h, status_rc, "a cell exceeded its capacity during generation");
  return NF_OK;
}

extern "C" int nf_get_rods(nf_handle* h, int inst, float* rods, uint32_t cap, uint32_t* n) {
  if (!h || inst < 0 || inst >= (int)h->params.size()) return fail(h, NF_E_INVALID, "bad instance");
  const uint32_t cnt = h->n_rods[inst];
  if (n) *n = cnt;
  if (!rods) return NF_OK;
  if (cap < cnt) return fail(h, NF_E_CAPACITY, "rod buffer too small");
  std::vector<float4> P(cnt), D(cnt);
  const uint32_t off = h->desc[inst].rod_off;
  if (cnt) {
    HIPCHK(h, hipMemcpy(P.data(), h->gb.bP + off, sizeof(float4) * cnt, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(D.data(), h->gb.bD + off, sizeof(float4) * cnt, hipMemcpyDeviceToHost));
  }
  for (uint32_t i = 0; i < cnt; ++i) {
    float* r = rods + 8 * (size_t)i;
    r[0] = P[i].x;
    r[1] = P[i].y;
    r[2] = P[i].z;
    r[3] = D[i].x;
    r[4] = D[i].y;
    r[5] = D[i].z;
    r[6] = 2.0f * P[i].w;
    r[7] = D[i].w;
  }
  return NF_OK;
}

extern "C" int nf_set_rods(nf_handle* h, int inst, const float* rods, uint32_t n) {
  if (!h || inst < 0 || inst >= (int)h->params.size()) return fail(h, NF_E_INVALID, "bad instance");
  if (n > h->desc[inst].n_target) return fail(h, NF_E_CAPACITY, "more rods than the instance target");
  if (n && !rods) return fail(h, NF_E_INVALID, "null rods");
  std::vector<float4> P(n), D(n);
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = rods + 8 * (size_t)i;
    if (!(r[6] > 0.0f) || !(r[7] > 0.0f)) return fail(h, NF_E_INVALID, "rod length and radius must be > 0");
    for (int a = 0; a < 3; ++a)
      if (!(r[a] >= 0.0f && r[a] < (float)h->params[inst].box[a]))
        return fail(h, NF_E_INVALID, "rod centre outside the box");
    P[i] = make_float4(r[0], r[1], r[2], 0.5f * r[6]);
    D[i] = make_float4(r[3], r[4], r[5], r[7]);
  }
  const uint32_t off = h->desc[inst].rod_off;
  if (n) {
    HIPCHK(h, hipMemcpy(h->gb.bP + off, P.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->gb.bD + off, D.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
  }
  h->n_rods[inst] = n;
  h->generated = false;
  h->cb.valid = false;
  h->cb.net_valid = false;
  h->cb.nd_valid = false;
  return NF_OK;
}


extern "C" int nf_get_params(nf_handle* h, int inst, nf_params* out) {
  if (!h || !out || inst < 0 || inst >= (int)h->params.size()) return fail(h, NF_E_INVALID, "bad instance");
  *out = h->params[inst];
  return NF_OK;
}

extern "C" int nf_agglomerate_centres(const nf_params* p, double* out, int32_t cap) {
  if (!p) return fail(nullptr, NF_E_INVALID, "null parameters");
  if (p->den_law != NF_DEN_AGGLOM) return 0;
  if (p->n_agg < 0) return fail(nullptr, NF_E_INVALID, "n_agg must be >= 0");
  if (out && cap < p->n_agg) return fail(nullptr, NF_E_CAPACITY, "buffer too small");
  if (out)
    for (int k = 0; k < p->n_agg; ++k) agg_centre(p->seed, p->stream, p->box, (uint32_t)k, &out[3 * k]);
  return p->n_agg;
}

// ===========================================================================
// diagnostics
// ===========================================================================
extern "C" int nf_selftest(int which, const double* in, double* out, uint64_t n) {
  if (which < 0 || which > 6 || (n && (!in || !out))) return fail(nullptr, NF_E_INVALID, "bad selftest args");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, NF_E_NODEVICE, "no HIP device");
  const uint64_t nin = which == 3 ? 2 * n : n;
  double *din = nullptr, *dout = nullptr;
  if (hipMalloc((void**)&din, sizeof(double) * (nin ? nin : 1)) != hipSuccess)
    return fail(nullptr, NF_E_NOMEM, "hipMalloc failed");
  if (hipMalloc((void**)&dout, sizeof(double) * (n ? n : 1)) != hipSuccess) {
    (void)hipFree(din);
    return fail(nullptr, NF_E_NOMEM, "hipMalloc failed");
  }
  hipError_t e = hipMemcpy(din, in, sizeof(double) * nin, hipMemcpyHostToDevice);
  if (e == hipSuccess && n) {
    launch_selftest(which, din, dout, n, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost);
  (void)hipFree(din);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(nullptr, NF_E_HIP, std::string("selftest: ") + hipGetErrorString(e));
  return NF_OK;
}

extern "C" int nf_selftest_pairs(const float* ra, const float* rb, const double* shift, uint64_t n, int contact,
                                 double sep_or_gap, double soft_frac, float ew, int32_t* screen, int32_t* decided,
                                 int32_t* within) {
  if (n && (!ra || !rb || !shift || !screen || !decided || !within))
    return fail(nullptr, NF_E_INVALID, "bad selftest_pairs args");
  if (n == 0) return NF_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, NF_E_NODEVICE, "no HIP device");
  const size_t fb = sizeof(float) * 8 * n, db = sizeof(double) * 3 * n, ib = sizeof(int32_t) * n;
  char* buf = nullptr;
  if (hipMalloc((void**)&buf, 2 * fb + db + 3 * ib) != hipSuccess) return fail(nullptr, NF_E_NOMEM, "hipMalloc failed");
  float* da = (float*)buf;
  float* dbr = (float*)(buf + fb);
  double* dsh = (double*)(buf + 2 * fb);
  int32_t* o = (int32_t*)(buf + 2 * fb + db);
  hipError_t e = hipMemcpy(da, ra, fb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dbr, rb, fb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dsh, shift, db, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_selftest_pairs(da, dbr, dsh, n, contact, sep_or_gap, soft_frac, ew, o, o + n, o + 2 * n, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(screen, o, ib, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(decided, o + n, ib, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(within, o + 2 * n, ib, hipMemcpyDeviceToHost);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(nullptr, NF_E_HIP, std::string("selftest_pairs: ") + hipGetErrorString(e));
  return NF_OK;
}
