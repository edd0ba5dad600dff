// nf_state.cpp - the reference's binary state file (output.bin,
// state_<vf>.bin): save_state_BIN / read_state_BIN, src/stateSR.cpp:54-134.
//
// Layout (native little-endian, no padding):
//   f32 physDim[3]   box edges
//   f32 edge[3]      cell edges
//   i32 ext[3]       cells per axis
//   f32 l, a         default length and radius
//   i32 ttlCNT       slots in the file
//   i32 numCNT       rods
//   f32 TOL
//   i16 occ[G]       rods per cell, G = ext0*ext1*ext2
//   i32 addr[G]      inclusive prefix of the cell capacities: cell i owns
//                    slots [addr[i-1], addr[i]) and fills the first occ[i]
//   f32 data[9][ttlCNT]  field-major x y z cx cy cz l a k, k = cell index
//                    (x + ext0*(y + ext1*z)); empty slots hold l = 0
//
// The writer repacks (capacity = occupancy, as simPrms::repack,
// src/simPrms.cpp:691-743, leaves it before save_state_BIN), so ttlCNT =
// numCNT.  The reader keeps every slot with l > TOL, the rule of
// simPrms::asmbl_cnt_data (src/simPrms.cpp:329-391), in slot order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "nf_host.h"

using nf::fail;

namespace {

constexpr size_t HEADER_BYTES = 3 * 4 + 3 * 4 + 3 * 4 + 4 + 4 + 4 + 4 + 4;  // 56
constexpr float STATE_TOL = 1e-6f;  // TOL default, src/simPrms.cpp:53-57

struct RawHeader {
  float box[3], edge[3];
  int32_t ext[3];
  float len, radius;
  int32_t slots, n;
  float tol;
};

template <class T>
void put(std::vector<char>& out, const T& v) {
  const char* p = reinterpret_cast<const char*>(&v);
  out.insert(out.end(), p, p + sizeof(T));
}

template <class T>
T get(const char* p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

RawHeader parse_header(const char* p) {
  RawHeader r;
  for (int d = 0; d < 3; ++d) r.box[d] = get<float>(p + 4 * d);
  for (int d = 0; d < 3; ++d) r.edge[d] = get<float>(p + 12 + 4 * d);
  for (int d = 0; d < 3; ++d) r.ext[d] = get<int32_t>(p + 24 + 4 * d);
  r.len = get<float>(p + 36);
  r.radius = get<float>(p + 40);
  r.slots = get<int32_t>(p + 44);
  r.n = get<int32_t>(p + 48);
  r.tol = get<float>(p + 52);
  return r;
}

// Reads the whole file and checks that its size matches the header.
int read_file(nf_handle* h, const char* path, std::vector<char>& buf, RawHeader& hd, uint64_t& G) {
  std::ifstream f(path, std::ios::in | std::ios::binary | std::ios::ate);
  if (!f.is_open()) return fail(h, NF_E_IO, std::string("cannot open state file ") + path);
  const std::streamoff size = f.tellg();
  if (size < (std::streamoff)HEADER_BYTES)
    return fail(h, NF_E_INVALID, std::string(path) + ": shorter than the state header");
  buf.resize((size_t)size);
  f.seekg(0);
  if (!f.read(buf.data(), size)) return fail(h, NF_E_IO, std::string("cannot read ") + path);
  hd = parse_header(buf.data());
  for (int d = 0; d < 3; ++d)
    if (hd.ext[d] < 1 || !(hd.box[d] > 0.0f))
      return fail(h, NF_E_INVALID, std::string(path) + ": bad box or cell extents in the header");
  G = (uint64_t)hd.ext[0] * (uint64_t)hd.ext[1] * (uint64_t)hd.ext[2];
  if (hd.slots < 0 || hd.n < 0 || G > (1ull << 30))
    return fail(h, NF_E_INVALID, std::string(path) + ": bad slot or rod count in the header");
  const uint64_t want = HEADER_BYTES + G * 2 + G * 4 + 9ull * 4 * (uint64_t)hd.slots;
  if ((uint64_t)size != want)
    return fail(h, NF_E_INVALID, std::string(path) + ": size " + std::to_string(size) +
                                     " bytes does not match the header (" + std::to_string(want) + ")");
  return NF_OK;
}

}  // namespace

extern "C" int nf_save_state(nf_handle* h, const char* path) {
  if (!h || !path) return fail(h, NF_E_INVALID, "null argument");
  const nf_params& p = h->params[h->post_inst];
  const nf::InstDesc& d = h->desc[h->post_inst];
  const uint32_t n = h->n_rods[h->post_inst];
  std::vector<float> rods((size_t)n * 8);
  uint32_t got = 0;
  int rc = nf_get_rods(h, h->post_inst, rods.data(), n, &got);
  if (rc != NF_OK) return rc;
  const int32_t ext[3] = {d.ncell[0], d.ncell[1], d.ncell[2]};
  const uint64_t G = (uint64_t)ext[0] * ext[1] * ext[2];
  // cell of every rod, as the reference's key: floor(x * ext / L) per axis
  std::vector<uint32_t> key(got), occ(G, 0);
  for (uint32_t i = 0; i < got; ++i) {
    int c[3];
    for (int a = 0; a < 3; ++a) {
      double v = std::floor((double)rods[8 * (size_t)i + a] * (double)ext[a] / p.box[a]);
      if (v < 0.0) v = 0.0;
      if (v > ext[a] - 1) v = ext[a] - 1;
      c[a] = (int)v;
    }
    key[i] = (uint32_t)(c[0] + ext[0] * (c[1] + ext[1] * c[2]));
    if (++occ[key[i]] > 32767u)
      return fail(h, NF_E_INVALID, "a cell holds more than 32767 rods: the state format's occupancy is 16-bit");
  }
  std::vector<uint32_t> start(G + 1, 0);
  for (uint64_t c = 0; c < G; ++c) start[c + 1] = start[c] + occ[c];
  std::vector<uint32_t> slot_of(got), fillc(G, 0);
  for (uint32_t i = 0; i < got; ++i) slot_of[i] = start[key[i]] + fillc[key[i]]++;
  std::vector<char> out;
  out.reserve(HEADER_BYTES + G * 6 + 36ull * got);
  for (int a = 0; a < 3; ++a) put(out, (float)p.box[a]);
  for (int a = 0; a < 3; ++a) put(out, (float)(p.box[a] / ext[a]));
  for (int a = 0; a < 3; ++a) put(out, ext[a]);
  put(out, (float)p.len);
  put(out, (float)p.radius);
  put(out, (int32_t)got);
  put(out, (int32_t)got);
  put(out, STATE_TOL);
  for (uint64_t c = 0; c < G; ++c) put(out, (int16_t)occ[c]);
  for (uint64_t c = 0; c < G; ++c) put(out, (int32_t)start[c + 1]);
  std::vector<float> field(got);
  for (int fi = 0; fi < 9; ++fi) {
    for (uint32_t i = 0; i < got; ++i)
      field[slot_of[i]] = fi < 8 ? rods[8 * (size_t)i + fi] : (float)key[i];
    const char* q = reinterpret_cast<const char*>(field.data());
    out.insert(out.end(), q, q + sizeof(float) * got);
  }
  std::ofstream f(path, std::ios::out | std::ios::binary | std::ios::trunc);
  if (!f.is_open()) return fail(h, NF_E_IO, std::string("cannot write ") + path);
  f.write(out.data(), (std::streamsize)out.size());
  if (!f.good()) return fail(h, NF_E_IO, std::string("write failed: ") + path);
  return NF_OK;
}

extern "C" int nf_read_state_header(const char* path, nf_state_header* out) {
  if (!path || !out) return fail(nullptr, NF_E_INVALID, "null argument");
  std::vector<char> buf;
  RawHeader hd;
  uint64_t G = 0;
  const int rc = read_file(nullptr, path, buf, hd, G);
  if (rc != NF_OK) return rc;
  for (int d = 0; d < 3; ++d) {
    out->box[d] = hd.box[d];
    out->edge[d] = hd.edge[d];
    out->ext[d] = hd.ext[d];
  }
  out->len = hd.len;
  out->radius = hd.radius;
  out->slots = (uint32_t)hd.slots;
  out->n_rods = (uint32_t)hd.n;
  out->tol = hd.tol;
  return NF_OK;
}

extern "C" int nf_load_state(nf_handle* h, const char* path, uint32_t* n_rods) {
  if (!h || !path) return fail(h, NF_E_INVALID, "null argument");
  std::vector<char> buf;
  RawHeader hd;
  uint64_t G = 0;
  int rc = read_file(h, path, buf, hd, G);
  if (rc != NF_OK) return rc;
  const nf_params& p = h->params[h->post_inst];
  for (int d = 0; d < 3; ++d)
    if (hd.box[d] != (float)p.box[d])
      return fail(h, NF_E_INVALID, std::string(path) + ": box " + std::to_string(hd.box[d]) + " on axis " +
                                       std::to_string(d) + " differs from the handle's " +
                                       std::to_string(p.box[d]));
  const size_t data0 = HEADER_BYTES + G * 6;
  const size_t S = (size_t)hd.slots;
  auto fieldp = [&](int fi, size_t i) { return get<float>(buf.data() + data0 + 4 * (fi * S + i)); };
  std::vector<float> rods;
  rods.reserve((size_t)hd.n * 8);
  for (size_t i = 0; i < S; ++i) {
    if (!(fieldp(6, i) > hd.tol)) continue;
    for (int fi = 0; fi < 8; ++fi) rods.push_back(fieldp(fi, i));
  }
  const uint32_t n = (uint32_t)(rods.size() / 8);
  rc = nf_set_rods(h, h->post_inst, n ? rods.data() : nullptr, n);
  if (rc != NF_OK) return rc;
  if (n_rods) *n_rods = n;
  return NF_OK;
}

// ===========================================================================
// expand (simPrms::expand, src/simPrms.cpp:744-852): the reference copies,
// for each new boundary cell, the rods of the old cell NeiOrder cells away
// across the periodic boundary, on the host.  Here the copies are chosen by
// centre on the device (flags, scans for their id-ordered ranks, one scatter
// into the new handle's rod buffers).
// ===========================================================================
extern "C" int nf_expand(nf_handle* h, int dir, nf_handle** out) {
  using namespace nf;
  if (!out) return fail(h, NF_E_INVALID, "null output handle");
  *out = nullptr;
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  if (dir != 1 && dir != 2 && dir != 4) return fail(h, NF_E_INVALID, "dir must be 1 (x), 2 (y) or 4 (z)");
  const int k = dir >> 1;
  const nf_params& p = h->params[h->post_inst];
  const nf::InstDesc& d = h->desc[h->post_inst];
  const uint32_t n = h->n_rods[h->post_inst];
  const float4* P = h->gb.bP + d.rod_off;
  const float4* D = h->gb.bD + d.rod_off;
  float hlmax = 0.0f, amax = 0.0f;
  int rc = contact_bounds(h, n, P, D, &hlmax, &amax);
  if (rc != NF_OK) return rc;
  const double L = p.box[k];
  const double E = L / (double)d.ncell[k];
  const double reach = 2.0 * (double)hlmax + p.contact_gap;
  const double B = std::ceil(reach / E) * E;
  if (!(B <= L)) return fail(h, NF_E_INVALID, "expand: the buffer (rod length + contact gap) is wider than the box");
  hipStream_t s = h->stream;
  uint32_t *below = nullptr, *above = nullptr, *rb = nullptr, *ra = nullptr, *tmp = nullptr;
  // flags per rod; their exclusive scans (ranks) plus the total at [n]
  if (rc == NF_OK) rc = dalloc(h, &below, std::max<uint64_t>(n, 1));
  if (rc == NF_OK) rc = dalloc(h, &above, std::max<uint64_t>(n, 1));
  if (rc == NF_OK) rc = dalloc(h, &rb, (uint64_t)n + 1);
  if (rc == NF_OK) rc = dalloc(h, &ra, (uint64_t)n + 1);
  if (rc == NF_OK) rc = dalloc(h, &tmp, scan_tmp_words(n));
  auto release = [&]() {
    dfree(h, below);
    dfree(h, above);
    dfree(h, rb);
    dfree(h, ra);
    dfree(h, tmp);
  };
  if (rc != NF_OK) {
    release();
    return rc;
  }
  launch_expand_flags(P, n, k, L, B, below, above, s);
  scan_u32(below, rb, n, tmp, s);
  scan_u32(above, ra, n, tmp, s);
  uint32_t nb = 0, na = 0;
  hipError_t e = hipMemcpyAsync(&nb, rb + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&na, ra + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = sync_stream(s, "expand count");
  if (e != hipSuccess) {
    release();
    return fail(h, NF_E_HIP, std::string("expand: ") + hipGetErrorString(e));
  }
  const uint64_t total = (uint64_t)n + nb + na;
  if (total > 0xFFFFFFF0ull) {
    release();
    return fail(h, NF_E_INVALID, "expand: too many rods");
  }
  nf_params q = p;
  q.box[k] = L + 2.0 * B;
  q.n_rods = (uint32_t)total;
  q.vol_frac = 0.0;
  nf_handle* g = nullptr;
  rc = nf_create(&q, &g);
  if (rc != NF_OK) {
    release();
    return fail(h, rc, std::string("expand: ") + g_last_error);
  }
  g->opt_contact_prec = h->opt_contact_prec;
  g->opt_contact_engine = h->opt_contact_engine;
  g->opt_footprint_fine = h->opt_footprint_fine;
  g->opt_footprint_occ = h->opt_footprint_occ;
  launch_expand_scatter(P, D, n, k, L, B, below, rb, above, ra, nb, g->gb.bP + g->desc[0].rod_off,
                        g->gb.bD + g->desc[0].rod_off, s);
  e = hipGetLastError();
  if (e == hipSuccess) e = sync_stream(s, "expand scatter");
  release();
  if (e != hipSuccess) {
    nf_destroy(g);
    return fail(h, NF_E_HIP, std::string("expand: ") + hipGetErrorString(e));
  }
  g->n_rods[0] = (uint32_t)total;
  g->generated = false;
  trace("expand: axis %d, buffer %.6f um, %u rods + %u below + %u above", k, B, n, nb, na);
  *out = g;
  return NF_OK;
}
