// Host build of the library's pair predicates (test infrastructure): lets
// the CPU test suite fuzz the FP32 error band against the FP64 oracle.
#include "../../nanofiller_amd/csrc/nf_common.h"
#include "../../nanofiller_amd/csrc/nf_seg.h"

#include <algorithm>
#include <vector>

extern "C" {

// rods: 8 floats each (x y z cx cy cz l a); shifts: 3 doubles per pair (the
// periodic image of rod b).  out_dec: -1 undecided, 0/1 cut result;
// out_within: reach result when decided; out_d2: FP32 squared distance.
void nfh_pair_f32_batch(int n, const float* ra, const float* rb, const double* sh, int contact,
                        double sep_or_gap, double soft_frac, float ew, int* out_dec,
                        int* out_within, float* out_d2) {
  for (int i = 0; i < n; ++i) {
    const float* a = ra + 8 * i;
    const float* b = rb + 8 * i;
    const float4 P1 = make_float4(a[0], a[1], a[2], 0.5f * a[6]);
    const float4 D1 = make_float4(a[3], a[4], a[5], a[7]);
    const float4 P2 = make_float4(b[0], b[1], b[2], 0.5f * b[6]);
    const float4 D2 = make_float4(b[3], b[4], b[5], b[7]);
    const float R32 = contact ? ((D1.w + D2.w) + (float)sep_or_gap)
                              : ((float)soft_frac * (D1.w + D2.w) + (float)sep_or_gap);
    const nf::F32Pair r = nf::pair_f32(P1, D1, P2, D2, (float)sh[3 * i], (float)sh[3 * i + 1],
                                       (float)sh[3 * i + 2], R32, ew, !contact);
    out_dec[i] = r.decided;
    out_within[i] = r.within;
    out_d2[i] = r.d2;
  }
}

void nfh_pair_screen_batch(int n, const float* ra, const float* rb, const double* sh, int contact,
                           double sep_or_gap, double soft_frac, float ew, int* out) {
  for (int i = 0; i < n; ++i) {
    const float* a = ra + 8 * i;
    const float* b = rb + 8 * i;
    const float4 P1 = make_float4(a[0], a[1], a[2], 0.5f * a[6]);
    const float4 D1 = make_float4(a[3], a[4], a[5], a[7]);
    const float4 P2 = make_float4(b[0], b[1], b[2], 0.5f * b[6]);
    const float4 D2 = make_float4(b[3], b[4], b[5], b[7]);
    const float R32 = contact ? ((D1.w + D2.w) + (float)sep_or_gap)
                              : ((float)soft_frac * (D1.w + D2.w) + (float)sep_or_gap);
    out[i] = nf::pair_f32_screen(P1, D1, P2, D2, (float)sh[3 * i], (float)sh[3 * i + 1],
                                 (float)sh[3 * i + 2], R32, ew);
  }
}

double nfh_sqrt_cr(double x) { return nf::sqrt_cr(x); }

// FP32 bound on a coordinate difference, as the contact pass computes it.
static double contact_ew(const double box[3], double hl_max, double a_max, double gap) {
  const double reach = (hl_max + hl_max) * nf::REACH_PAD + ((a_max + a_max) + gap);
  const double L = std::max(box[0], std::max(box[1], box[2]));
  return 5.9604644775390625e-8 * (L + reach) * 1.01;
}

static void rod_f4(const float* r, float4* P, float4* D) {
  *P = make_float4(r[0], r[1], r[2], 0.5f * r[6]);
  *D = make_float4(r[3], r[4], r[5], r[7]);
}

// Fine mask (seg_fine_mask) and crossing test (seg_clip) of one rod against
// the home box [lo, lo + f e) with footprint radius r; returns the mask,
// *cross the crossing result.
uint64_t nfh_fine_mask(const float* rod, const float lo[3], const float e[3], int f, float r,
                       int* cross) {
  float4 P, D;
  rod_f4(rod, &P, &D);
  const float p[3] = {P.x, P.y, P.z};
  const float c[3] = {D.x, D.y, D.z};
  float hi[3];
  for (int k = 0; k < 3; ++k) hi[k] = lo[k] + (float)f * e[k];
  float s0, s1;
  *cross = nf::seg_clip(p, c, P.w, lo, hi, r, &s0, &s1) ? 1 : 0;
  const float ie[3] = {1.0f / e[0], 1.0f / e[1], 1.0f / e[2]};
  return nf::seg_fine_mask(p, c, P.w, lo, e, ie, f, r);
}

// Near mask (seg_near_mask) of a rod binned in cell cc of a grid of cell
// edges h: the 27-bit mask, or SEG_LONG.
uint32_t nfh_near_mask(const float* rod, const int cc[3], const float h[3], float r, float mx) {
  float4 P, D;
  rod_f4(rod, &P, &D);
  const float p[3] = {P.x, P.y, P.z};
  const float c[3] = {D.x, D.y, D.z};
  return nf::seg_near_mask(p, c, P.w, cc, h, r, mx);
}

// Host emulation of the footprint contact pass of nf_contacts on home cells
// of about `edge` with f fine cells per axis: every rod's capsule (radius
// from seg_grid) is tested against every home cell it can reach, each home
// cell's survivors get their fine masks, and every pair sharing a fine cell
// goes through seg_test with the home cell as the claim.  row > 1 finds a
// cell's crossing rods as the kernel does, per run of `row` cells along x
// (seg_row_cells); row 1 clips every rod against every cell.  Fills I, J,
// image, gap64 (unsorted, up to cap) and returns the number of reported
// contacts; *max_surv the largest survivor count of a home cell.
int64_t nfh_footprint_contacts(const float* rods, uint32_t n, const double box[3], double edge, int f,
                               double gap, int prec, int row, int near, uint32_t* I, uint32_t* J, int32_t* img,
                               double* gap64, int64_t cap, uint64_t* max_surv) {
  std::vector<float4> P(n), D(n);
  float hl_max = 0.0f, a_max = 0.0f;
  for (uint32_t i = 0; i < n; ++i) {
    rod_f4(rods + 8 * (size_t)i, &P[i], &D[i]);
    hl_max = std::max(hl_max, P[i].w);
    a_max = std::max(a_max, D[i].w);
  }
  const double ew = contact_ew(box, hl_max, a_max, gap);
  const nf::SegGrid g = nf::seg_grid(box, edge, a_max, gap, ew);
  struct Surv {
    uint32_t rod, meta;
  };
  std::vector<std::vector<Surv>> cells(g.ncells);
  for (uint32_t i = 0; i < n; ++i) {
    const float hl = P[i].w, r = g.r;
    int lo_u[3], hi_u[3];
    const float pp[3] = {P[i].x, P[i].y, P[i].z};
    for (int k = 0; k < 3; ++k) {
      lo_u[k] = (int)std::floor((pp[k] - hl - r) * g.inv_h[k]) - 1;
      hi_u[k] = (int)std::floor((pp[k] + hl + r) * g.inv_h[k]) + 1;
    }
    if (near) {
      // the kernel's near masks: the rod's binned cell (FP64 product,
      // truncated, clamped), the cells of its 3 x 3 x 3 block it crosses;
      // a rod that may leave the block is tested per work unit below
      int cc[3];
      for (int k = 0; k < 3; ++k) {
        const int q = (int)((double)pp[k] * ((double)g.nc[k] / box[k]));
        cc[k] = q < 0 ? 0 : (q >= g.nc[k] ? g.nc[k] - 1 : q);
      }
      const float c[3] = {D[i].x, D[i].y, D[i].z};
      const uint32_t m = nf::seg_near_mask(pp, c, hl, cc, g.h, r, 4.0f * (float)ew);
      if (m != nf::SEG_LONG) {
        for (int b = 0; b < 27; ++b) {
          if (((m >> b) & 1u) == 0u) continue;
          const int u[3] = {cc[0] + b % 3 - 1, cc[1] + (b / 3) % 3 - 1, cc[2] + b / 9 - 1};
          int w[3], t[3];
          for (int k = 0; k < 3; ++k) {
            const int qd = nf::floor_div(u[k], g.nc[k]);
            w[k] = u[k] - qd * g.nc[k];
            t[k] = -qd;
          }
          cells[(size_t)w[0] + (size_t)g.nc[0] * ((size_t)w[1] + (size_t)g.nc[1] * (size_t)w[2])].push_back(
              {i, nf::seg_pack(t[0], t[1], t[2])});
        }
        continue;
      }
    }
    if (row > 1) {
      // the kernel's work units: runs of `row` cells along x, one crossing
      // test per run (seg_row_cells) giving the run's cells the rod crosses
      const float mx = 4.0f * (float)ew;
      const float c[3] = {D[i].x, D[i].y, D[i].z};
      for (int u1 = lo_u[1]; u1 <= hi_u[1]; ++u1)
        for (int u2 = lo_u[2]; u2 <= hi_u[2]; ++u2) {
          int last_q = 0, last_seg = -1;
          for (int u0 = lo_u[0]; u0 <= hi_u[0]; ++u0) {
            const int u[3] = {u0, u1, u2};
            int wc[3], t[3];
            float q[3];
            for (int k = 0; k < 3; ++k) {
              const int qd = nf::floor_div(u[k], g.nc[k]);
              wc[k] = u[k] - qd * g.nc[k];
              t[k] = -qd;
              q[k] = pp[k] + (float)t[k] * (float)box[k];
            }
            const int seg = wc[0] / row;
            if (seg == last_seg && -t[0] == last_q) continue;
            last_seg = seg;
            last_q = -t[0];
            const int x0 = seg * row, w = std::min(row, g.nc[0] - x0);
            const float lo[3] = {(float)x0 * g.h[0], (float)wc[1] * g.h[1], (float)wc[2] * g.h[2]};
            const float hi[3] = {(float)(x0 + w) * g.h[0], (float)(wc[1] + 1) * g.h[1], (float)(wc[2] + 1) * g.h[2]};
            const uint32_t bits = nf::seg_row_cells(q, c, hl, lo, hi, x0, w, g.h[0], r, mx);
            for (int b = 0; b < w; ++b)
              if ((bits >> b) & 1u)
                cells[(size_t)(x0 + b) + (size_t)g.nc[0] * ((size_t)wc[1] + (size_t)g.nc[1] * (size_t)wc[2])]
                    .push_back({i, nf::seg_pack(t[0], t[1], t[2])});
          }
        }
      continue;
    }
    for (int u0 = lo_u[0]; u0 <= hi_u[0]; ++u0)
      for (int u1 = lo_u[1]; u1 <= hi_u[1]; ++u1)
        for (int u2 = lo_u[2]; u2 <= hi_u[2]; ++u2) {
          const int u[3] = {u0, u1, u2};
          int w[3], t[3];
          float lo[3], hi[3], q[3];
          for (int k = 0; k < 3; ++k) {
            const int qd = nf::floor_div(u[k], g.nc[k]);
            w[k] = u[k] - qd * g.nc[k];
            t[k] = -qd;  // the rod moved by t box edges meets the home cell w
            lo[k] = (float)w[k] * g.h[k];
            hi[k] = (float)(w[k] + 1) * g.h[k];
            q[k] = pp[k] + (float)t[k] * (float)box[k];
          }
          const float c[3] = {D[i].x, D[i].y, D[i].z};
          float s0, s1;
          if (!nf::seg_clip(q, c, hl, lo, hi, r, &s0, &s1)) continue;
          cells[(size_t)w[0] + (size_t)g.nc[0] * ((size_t)w[1] + (size_t)g.nc[1] * (size_t)w[2])].push_back(
              {i, nf::seg_pack(t[0], t[1], t[2])});
        }
  }
  nf::SegCtx x{};
  for (int k = 0; k < 3; ++k) {
    x.box[k] = box[k];
    x.boxf[k] = (float)box[k];
  }
  x.gap64 = gap;
  x.gap32 = (float)gap;
  x.ew = (float)ew;
  int64_t out = 0;
  uint32_t tests = 0, rech = 0;
  double maxdd = 0.0;
  *max_surv = 0;
  for (uint32_t cell = 0; cell < g.ncells; ++cell) {
    const std::vector<Surv>& sv = cells[cell];
    if (sv.size() > *max_surv) *max_surv = sv.size();
    const int hc[3] = {(int)(cell % (uint32_t)g.nc[0]), (int)((cell / (uint32_t)g.nc[0]) % (uint32_t)g.nc[1]),
                       (int)(cell / ((uint32_t)g.nc[0] * (uint32_t)g.nc[1]))};
    float lo[3], e[3], ie[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = (float)hc[k] * g.h[k];
      e[k] = g.h[k] / (float)f;
      ie[k] = 1.0f / e[k];
    }
    std::vector<uint64_t> fm(sv.size());
    for (size_t a = 0; a < sv.size(); ++a) {
      const uint32_t i = sv[a].rod;
      const float p[3] = {P[i].x + (float)nf::seg_img(sv[a].meta, 0) * (float)box[0],
                          P[i].y + (float)nf::seg_img(sv[a].meta, 1) * (float)box[1],
                          P[i].z + (float)nf::seg_img(sv[a].meta, 2) * (float)box[2]};
      const float c[3] = {D[i].x, D[i].y, D[i].z};
      fm[a] = nf::seg_fine_mask(p, c, P[i].w, lo, e, ie, f, g.r);
    }
    for (size_t a = 0; a < sv.size(); ++a)
      for (size_t b = a + 1; b < sv.size(); ++b) {
        if ((fm[a] & fm[b]) == 0ull || sv[a].rod == sv[b].rod) continue;
        const uint32_t ra = sv[a].rod, rb = sv[b].rod;
        nf::SegHit hit;
        const bool rep =
            prec == NF_PREC_FP64
                ? nf::seg_test<NF_PREC_FP64>(x, g, ra, P[ra], D[ra], sv[a].meta, rb, P[rb], D[rb], sv[b].meta,
                                             cell, tests, rech, maxdd, &hit)
                : nf::seg_test<NF_PREC_FP32_RECHECK>(x, g, ra, P[ra], D[ra], sv[a].meta, rb, P[rb], D[rb],
                                                     sv[b].meta, cell, tests, rech, maxdd, &hit);
        if (!rep) continue;
        if (out < cap) {
          I[out] = hit.I;
          J[out] = hit.J;
          img[out] = hit.img;
          gap64[out] = hit.gap;
        }
        ++out;
      }
  }
  return out;
}

}  // extern "C"
