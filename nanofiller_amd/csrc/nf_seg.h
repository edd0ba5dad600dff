// nf_seg.h - the footprint contact pass, shared by the HIP kernels
// (nf_seg.hip) and host builds: the grid, the crossing test and fine-cell
// footprint of one rod in a home cell, and the test of one candidate pair
// with its reporting rule.
//
// Thin rods in contact come within a_I + a_J + gap of each other, far less
// than the reach of their centres (l + 2a + gap).  The midpoint of a
// contact's closest points lies within r = (a_I + a_J + gap) / 2 of both
// axes, so both capsules of radius r (a_max + gap / 2 plus rounding
// margins) contain it: both rods cross the home cell that holds it, and both
// footprints on that cell's fine cells hold the midpoint's fine cell.  The
// candidate pairs of a home cell are its crossing rods that share a fine
// cell, and a contact is reported only by the home cell holding its
// midpoint, so exactly once.
#pragma once

#include <math.h>
#include <stdint.h>

#include "nf_common.h"

namespace nf {

// Cells of edge h = box / nc per axis.
struct SegGrid {
  int32_t nc[3];
  float h[3];         // cell edge (FP32)
  float inv_h[3];     // nc / box (FP32)
  double inv_h64[3];  // nc / box (FP64), for the reporting cell of a contact
  float r;            // footprint radius: a_max + gap / 2 with rounding margins
  uint32_t ncells;
};

// Grid of about `edge` per cell (at least one and at most 1024 cells per
// axis).  ew bounds the FP32 error of a coordinate difference (about 2^-24
// of box + reach); the radius margin covers the FP32 footprint arithmetic
// and the FP64 midpoint.
NF_HD SegGrid seg_grid(const double box[3], double edge, double a_max, double gap, double ew) {
  SegGrid g{};
  uint64_t nc = 1;
  for (int a = 0; a < 3; ++a) {
    double k = floor(box[a] / edge);
    if (k < 1.0) k = 1.0;
    if (k > 1024.0) k = 1024.0;
    g.nc[a] = (int32_t)k;
    g.h[a] = (float)(box[a] / k);
    g.inv_h[a] = (float)(k / box[a]);
    g.inv_h64[a] = k / box[a];
    nc *= (uint64_t)k;
  }
  g.ncells = (uint32_t)nc;
  g.r = (float)((a_max + 0.5 * gap) * 1.0001 + 64.0 * ew + 1e-7);
  return g;
}

NF_HD int seg_img(uint32_t m, int a) { return (int)(int8_t)((m >> (8 * a)) & 0xFFu); }
NF_HD uint32_t seg_pack(int x, int y, int z) {
  return ((uint32_t)(x & 0xFF)) | ((uint32_t)(y & 0xFF) << 8) | ((uint32_t)(z & 0xFF) << 16);
}

// floor(u / n) for n > 0
NF_HD int floor_div(int u, int n) {
  const int q = u / n;
  return (u % n != 0 && u < 0) ? q - 1 : q;
}

// Parameter interval [*s0, *s1] of the segment P + s c, s in [-hl, hl],
// inside the box [lo - r, hi + r] (max-norm dilation of [lo, hi] by r);
// false when empty.  FP32, each slab interval widened by a few ulps of its
// scale, so the interval contains the exact one.
NF_HD bool seg_clip(const float p[3], const float c[3], float hl, const float lo[3], const float hi[3],
                    float r, float* s0, float* s1) {
  float a = -hl, b = hl;
  for (int k = 0; k < 3; ++k) {
    const float l = lo[k] - r, u = hi[k] + r;
    if (c[k] == 0.0f) {
      if (p[k] < l || p[k] > u) return false;
      continue;
    }
    const float rc = recip(c[k]);
    float t0 = (l - p[k]) * rc, t1 = (u - p[k]) * rc;
    if (t0 > t1) {
      const float t = t0;
      t0 = t1;
      t1 = t;
    }
    const float ds = (fabsf(t0) + fabsf(t1) + hl) * 4.8e-7f;
    a = fmaxf(a, t0 - ds);
    b = fminf(b, t1 + ds);
  }
  *s0 = a;
  *s1 = b;
  return a <= b;
}

// Home cells of a run of w cells along x (the box [lo, hi]; cell i spans x
// in [(x0 + i) h0, (x0 + i + 1) h0]) that the segment's capsule of radius r
// crosses, as bits.  The cells share their y and z ranges and x(s) is
// linear, so the segment crosses cell i iff the x extent of its part inside
// the dilated run box meets the cell's x range dilated by r; mx covers the
// rounding of that extent (two FP32 operations on coordinates below box +
// reach).
NF_HD uint32_t seg_row_cells(const float p[3], const float c[3], float hl, const float lo[3], const float hi[3],
                             int x0, int w, float h0, float r, float mx) {
  float s0, s1;
  if (!seg_clip(p, c, hl, lo, hi, r, &s0, &s1)) return 0u;
  const float xa = p[0] + s0 * c[0], xb = p[0] + s1 * c[0];
  const float xl = fminf(xa, xb) - mx, xh = fmaxf(xa, xb) + mx;
  uint32_t cells = 0u;
  for (int i = 0; i < w; ++i) {
    const float cl = (float)(x0 + i) * h0 - r, ch = (float)(x0 + i + 1) * h0 + r;
    if (xh >= cl && xl <= ch) cells |= 1u << i;
  }
  return cells;
}

// Fine cells of one home box [lo, lo + f e) (f cells of edge e per axis,
// bit x + f (y + f z)) that the segment's capsule of radius r can touch: the
// segment clipped to the dilated box, then per slab of the dominant axis the
// rectangle of cells its range on the other axes (widened by r) meets.
// Cell indices are taken with ie = 1 / e (FP32): the product is within
// about an ulp of the quotient, a few 1e-7 of a cell, far inside the
// radius's margin (64 ew, at least 6e-6 um).
NF_HD uint64_t seg_fine_mask(const float p[3], const float c[3], float hl, const float lo[3],
                             const float e[3], const float ie[3], int f, float r) {
  float hi[3];
  for (int k = 0; k < 3; ++k) hi[k] = lo[k] + (float)f * e[k];
  float s0, s1;
  if (!seg_clip(p, c, hl, lo, hi, r, &s0, &s1)) return 0ull;
  if (f == 1) return 1ull;
  const float ax = fabsf(c[0]), ay = fabsf(c[1]), az = fabsf(c[2]);
  const int d0 = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  const int d1 = d0 == 0 ? 1 : 0;
  const int d2 = d0 == 2 ? 1 : 2;
  const float rc = recip(c[d0]);
  // bit x + f (y + f z) of fine cell (x, y, z): index strides of the slab
  // axis and the two others (d1 is x or y, d2 is y or z)
  const int t0 = d0 == 0 ? 1 : (d0 == 1 ? f : f * f);
  const int t1 = d1 == 0 ? 1 : f;
  const int t2 = d2 == 1 ? f : f * f;
  uint64_t m = 0ull;
  for (int i = 0; i < f; ++i) {
    const float xa = lo[d0] + (float)i * e[d0] - r, xb = lo[d0] + (float)(i + 1) * e[d0] + r;
    float sa = (xa - p[d0]) * rc, sb = (xb - p[d0]) * rc;
    if (sa > sb) {
      const float t = sa;
      sa = sb;
      sb = t;
    }
    const float ds = (fabsf(sa) + fabsf(sb) + hl) * 4.8e-7f;
    sa = fmaxf(sa - ds, s0);
    sb = fminf(sb + ds, s1);
    if (sa > sb) continue;
    const float ya = p[d1] + sa * c[d1], yb = p[d1] + sb * c[d1];
    const float za = p[d2] + sa * c[d2], zb = p[d2] + sb * c[d2];
    int j0 = (int)floorf((fminf(ya, yb) - r - lo[d1]) * ie[d1]);
    int j1 = (int)floorf((fmaxf(ya, yb) + r - lo[d1]) * ie[d1]);
    int k0 = (int)floorf((fminf(za, zb) - r - lo[d2]) * ie[d2]);
    int k1 = (int)floorf((fmaxf(za, zb) + r - lo[d2]) * ie[d2]);
    j0 = j0 < 0 ? 0 : j0;
    k0 = k0 < 0 ? 0 : k0;
    j1 = j1 > f - 1 ? f - 1 : j1;
    k1 = k1 > f - 1 ? f - 1 : k1;
    // the rectangle [j0, j1] x [k0, k1] of slab i: a row of bits along d1,
    // shifted along d2
    uint64_t row = 0ull;
    for (int j = j0; j <= j1; ++j) row |= 1ull << (j * t1);
    for (int k = k0; k <= k1; ++k) m |= row << (i * t0 + k * t2);
  }
  return m;
}

// Near mask of a rod binned in cell cc: the cells of the 3 x 3 x 3 block
// around cc (bit (dx + 1) + 3 (dy + 1) + 9 (dz + 1), cell boundaries k h)
// that its capsule of radius r can touch, the fine mask of that block with
// f = 3 and e = h.  SEG_LONG instead when the capsule may leave the block
// (a rod longer than about a cell, or a centre outside its cell): such a
// rod is tested against each work unit (seg_row_cells).  mx covers the
// rounding of the capsule's extent.  Depends on the rod alone, so it is
// computed once per rod instead of once per work unit that examines it.
constexpr uint32_t SEG_LONG = 1u << 31;

NF_HD uint32_t seg_near_mask(const float p[3], const float c[3], float hl, const int cc[3], const float h[3],
                             float r, float mx) {
  float lo[3], e[3], ie[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = (float)(cc[k] - 1) * h[k];
    const float hi = (float)(cc[k] + 2) * h[k];
    const float ext = fabsf(c[k]) * hl + r + mx;
    if (!(p[k] - ext >= lo[k] && p[k] + ext <= hi)) return SEG_LONG;
    e[k] = h[k];
    ie[k] = 1.0f / h[k];
  }
  return (uint32_t)seg_fine_mask(p, c, hl, lo, e, ie, 3, r);
}

// Home cells of a run of w cells (bit i: the run's cell i) crossed by a rod
// with near mask m (not SEG_LONG) in the cell at offset (ox, oy, oz) from
// the run's first cell.
NF_HD uint32_t seg_near_cells(uint32_t m, int ox, int oy, int oz, int w) {
  if (oy < -1 || oy > 1 || oz < -1 || oz > 1) return 0u;
  // the row of the block at the run's y and z: dx = -1, 0, 1 of the rod's cell
  const uint32_t row = (m >> (3 * (1 - oy) + 9 * (1 - oz))) & 7u;
  // bit dx + 1 is the run's cell ox + dx
  const int sh = ox - 1;
  uint32_t cells;
  if (sh >= 0)
    cells = sh < 32 ? row << sh : 0u;
  else
    cells = -sh < 32 ? row >> -sh : 0u;
  return cells & ((1u << w) - 1u);
}

struct SegHit {
  uint32_t I, J;  // lower and higher rod id
  int32_t img;    // J's periodic image (signed bytes x, y, z)
  double gap;     // FP64 surface gap
};

struct SegCtx {
  double box[3];
  float boxf[3];
  double gap64;
  float gap32;
  float ew;
  int diag;
};

// One candidate pair: rod a with its entry's translate ma, rod b with mb,
// met in cell `cell` of the grid.  The capsule test of the contact pass
// (FP32 screen and band, FP64 for undecided pairs, or FP64 alone) decides;
// a contact is reported only when the cell is the one that holds the
// midpoint of the canonical FP64 closest points (lower id first, the image
// on the higher id), unwrapped in the lower id's frame.  tests counts pairs
// within centre reach, rech the FP64 rechecks, maxdd the largest
// |d32 - d64| of FP32-decided pairs (ctx.diag).
//
// Two stages, so that a kernel can run the cheap screen on every pair and
// the rest on dense batches of the pairs it passes: seg_screen returns 0
// (settled: no contact), 1 (the FP32 band decides) or 2 (FP64 decides:
// reach undecided by the screen, or the plain FP64 path); seg_full takes
// that code and decides and reports.
template <int PREC>
NF_HD int seg_screen(const SegCtx& x, const float4 Pa, const float4 Da, uint32_t ma, const float4 Pb,
                     const float4 Db, uint32_t mb, uint32_t& tests) {
  if (PREC == NF_PREC_FP64) return 2;
  const int tx = seg_img(mb, 0) - seg_img(ma, 0);
  const int ty = seg_img(mb, 1) - seg_img(ma, 1);
  const int tz = seg_img(mb, 2) - seg_img(ma, 2);
  const float fx = (float)tx * x.boxf[0], fy = (float)ty * x.boxf[1], fz = (float)tz * x.boxf[2];
  const float R32 = (Da.w + Db.w) + x.gap32;
  const int sc = pair_f32_screen(Pa, Da, Pb, Db, fx, fy, fz, R32, x.ew);
  if (sc == SCREEN_OUT) return 0;
  if (sc == SCREEN_UNDECIDED) return 2;
  ++tests;
  return sc == SCREEN_FULL ? 1 : 0;
}

template <int PREC>
NF_HD bool seg_full(const SegCtx& x, const SegGrid& g, uint32_t ra, const float4 Pa, const float4 Da,
                    uint32_t ma, uint32_t rb, const float4 Pb, const float4 Db, uint32_t mb, uint32_t cell,
                    int code, uint32_t& tests, uint32_t& rech, double& maxdd, SegHit* hit) {
  // b relative to a
  const int tx = seg_img(mb, 0) - seg_img(ma, 0);
  const int ty = seg_img(mb, 1) - seg_img(ma, 1);
  const int tz = seg_img(mb, 2) - seg_img(ma, 2);
  const bool swap = rb < ra;
  const float4 PI = swap ? Pb : Pa, DI = swap ? Db : Da;
  const float4 PJ = swap ? Pa : Pb, DJ = swap ? Da : Db;
  const int jx = swap ? -tx : tx, jy = swap ? -ty : ty, jz = swap ? -tz : tz;
  const double sx = (double)jx * x.box[0], sy = (double)jy * x.box[1], sz = (double)jz * x.box[2];
  const double R64 = ((double)DI.w + (double)DJ.w) + x.gap64;
  bool cut;
  if (code == 2) {
    if (PREC != NF_PREC_FP64) ++rech;
    bool within;
    const double d2 = pair_d2_f64(PI, DI, PJ, DJ, sx, sy, sz, R64, &within);
    if (!within) return false;
    ++tests;
    cut = d2 <= R64 * R64;
  } else {
    const float fx = (float)tx * x.boxf[0], fy = (float)ty * x.boxf[1], fz = (float)tz * x.boxf[2];
    const float R32 = (Da.w + Db.w) + x.gap32;
    const F32Pair r = pair_f32(Pa, Da, Pb, Db, fx, fy, fz, R32, x.ew, false);
    if (r.decided < 0) {
      ++rech;
      bool w64;
      const double d2 = pair_d2_f64(PI, DI, PJ, DJ, sx, sy, sz, R64, &w64);
      cut = d2 <= R64 * R64;
    } else {
      cut = r.decided != 0;
      if (x.diag) {
        bool w2;
        const double e64 = pair_d2_f64(PI, DI, PJ, DJ, sx, sy, sz, R64, &w2);
        const double dd = fabs(sqrt((double)r.d2) - sqrt(e64));
        maxdd = dd > maxdd ? dd : maxdd;
      }
    }
  }
  if (!cut) return false;
  // the midpoint of the canonical closest points, in the lower id's frame
  bool w3;
  double s, t;
  const double d2 = pair_closest_f64(PI, DI, PJ, DJ, sx, sy, sz, R64, &w3, &s, &t);
  const double hI = (double)PI.w, hJ = (double)PJ.w;
  const double m[3] = {
      (((double)PI.x + s * (hI * (double)DI.x)) + (((double)PJ.x + sx) + t * (hJ * (double)DJ.x))) * 0.5,
      (((double)PI.y + s * (hI * (double)DI.y)) + (((double)PJ.y + sy) + t * (hJ * (double)DJ.y))) * 0.5,
      (((double)PI.z + s * (hI * (double)DI.z)) + (((double)PJ.z + sz) + t * (hJ * (double)DJ.z))) * 0.5};
  // this entry pair's cell, unwrapped in the lower id's frame
  const uint32_t mI = swap ? mb : ma;
  const uint32_t cw[3] = {cell % (uint32_t)g.nc[0], (cell / (uint32_t)g.nc[0]) % (uint32_t)g.nc[1],
                          cell / ((uint32_t)g.nc[0] * (uint32_t)g.nc[1])};
  for (int k = 0; k < 3; ++k) {
    const int u = (int)floor(m[k] * g.inv_h64[k]);
    if (u != (int)cw[k] - seg_img(mI, k) * g.nc[k]) return false;
  }
  hit->I = swap ? rb : ra;
  hit->J = swap ? ra : rb;
  hit->img = (int32_t)seg_pack(jx, jy, jz);
  hit->gap = sqrt_cr(d2) - ((double)DI.w + (double)DJ.w);
  return true;
}

template <int PREC>
NF_HD bool seg_test(const SegCtx& x, const SegGrid& g, uint32_t ra, const float4 Pa, const float4 Da,
                    uint32_t ma, uint32_t rb, const float4 Pb, const float4 Db, uint32_t mb, uint32_t cell,
                    uint32_t& tests, uint32_t& rech, double& maxdd, SegHit* hit) {
  const int code = seg_screen<PREC>(x, Pa, Da, ma, Pb, Db, mb, tests);
  if (code == 0) return false;
  return seg_full<PREC>(x, g, ra, Pa, Da, ma, rb, Pb, Db, mb, cell, code, tests, rech, maxdd, hit);
}

}  // namespace nf
