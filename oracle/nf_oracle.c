/*
 * nf_oracle.c - FP64 CPU oracle (test infrastructure; see nf_oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math): every
 * double expression below is evaluated exactly as written, one IEEE rounding
 * per operation, which is what makes the comparison with the GPU bit-exact.
 *
 * Pinning (tests/test_oracle.py): Philox known-answer vectors, exp/log
 * against libm, capsule distance against closed forms and brute-force
 * sampling, generation against brute-force all-pairs overlap checks and the
 * round-rule invariants, contacts against an O(N^2) sweep, histograms
 * against numpy recounts, components against BFS, CSR against Laplacian
 * invariants and hand-built series/parallel networks.
 */
#include "nf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================== */
/* Philox4x32-10                                                          */
/* ===================================================================== */
/* Replaces curand XORWOW seeded from clock() (src/simPrms.cpp:415, 853-866;
 * inc/cuda_CNTgenerate.h:200): a counter-based generator lets the CPU and
 * the GPU draw the same candidate stream without sharing state. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

double orc_u01(uint32_t a, uint32_t b) {
  uint64_t m = ((uint64_t)(a >> 12) << 32) | (uint64_t)b; /* 52 bits */
  return ((double)m + 0.5) * 0x1p-52;
}

static void blk(const nf_params* p, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                uint32_t o[4]) {
  uint32_t ctr[4] = {c0, c1, c2, c3};
  uint32_t key[2] = {p->seed, p->stream};
  orc_philox4x32_10(ctr, key, o);
}

/* ===================================================================== */
/* exp / log from basic operations                                         */
/* ===================================================================== */
/* The reference samples with device erf/erfinv/cosf (inc/cuda_CNTgenerate.h:
 * 262-275) which no CPU reproduces.  These use Cody-Waite range reduction
 * and Taylor / atanh series evaluated with +,-,*,/ only, so the CPU and the
 * GPU produce identical bits. */
static const double LN2_HI = 6.93147180369123816490e-01;
static const double LN2_LO = 1.90821492927058770002e-10;
static const double INV_LN2 = 1.44269504088896338700e+00;

static double pow2i(int k) { /* 2^k for -1022 <= k <= 1023 */
  union { uint64_t u; double d; } v;
  v.u = (uint64_t)(k + 1023) << 52;
  return v.d;
}

double orc_exp(double x) {
  if (x != x) return x;
  if (x > 709.0) return INFINITY;
  if (x < -745.0) return 0.0;
  double kd = floor(x * INV_LN2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * LN2_HI) - kd * LN2_LO;
  double p = 1.0 / 6227020800.0; /* 1/13! */
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  if (k > 1023) return (p * pow2i(1023)) * pow2i(k - 1023);
  if (k < -1022) return (p * pow2i(k + 600)) * pow2i(-600);
  return p * pow2i(k);
}

double orc_log(double x) {
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return x;
  int e = 0;
  if (x < 0x1p-1022) { /* subnormal */
    x = x * 0x1p54;
    e = -54;
  }
  union { uint64_t u; double d; } v;
  v.d = x;
  e += (int)((v.u >> 52) & 0x7FF) - 1023;
  v.u = (v.u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull; /* m in [1,2) */
  double m = v.d;
  if (m > 1.41421356237309504880) {
    m = m * 0.5;
    e += 1;
  }
  double f = (m - 1.0) / (m + 1.0);
  double f2 = f * f;
  double s = 1.0 / 21.0;
  s = s * f2 + 1.0 / 19.0;
  s = s * f2 + 1.0 / 17.0;
  s = s * f2 + 1.0 / 15.0;
  s = s * f2 + 1.0 / 13.0;
  s = s * f2 + 1.0 / 11.0;
  s = s * f2 + 1.0 / 9.0;
  s = s * f2 + 1.0 / 7.0;
  s = s * f2 + 1.0 / 5.0;
  s = s * f2 + 1.0 / 3.0;
  s = s * f2 + 1.0;
  double lm = (2.0 * f) * s;
  double ed = (double)e;
  return ed * LN2_HI + (lm + ed * LN2_LO);
}

/* ===================================================================== */
/* Target distributions                                                    */
/* ===================================================================== */
/* Agglomerate centres: Philox counter (k, 0, d, 1).  Stands in for the
 * density map the reference reads from file (src/IO.cpp:11-36) or builds
 * with the Fortran agglomerate tool (tools/Cond_map_gen/src/Mod_Aglomer.f90:
 * 105-130, which combines agglomerates with MAX as done below). */
void orc_agg_centres(const nf_params* p, double* out) {
  for (int k = 0; k < p->n_agg; ++k) {
    uint32_t o[4], q[4];
    blk(p, (uint32_t)k, 0u, 0u, 1u, o);
    blk(p, (uint32_t)k, 0u, 1u, 1u, q);
    out[3 * k + 0] = orc_u01(o[0], o[1]) * p->box[0];
    out[3 * k + 1] = orc_u01(o[2], o[3]) * p->box[1];
    out[3 * k + 2] = orc_u01(q[0], q[1]) * p->box[2];
  }
}

void orc_vmf_basis(const double mu[3], double e1[3], double e2[3]) {
  double t[3] = {0.0, 0.0, 0.0};
  double ax = fabs(mu[0]), ay = fabs(mu[1]), az = fabs(mu[2]);
  if (ax <= ay && ax <= az)
    t[0] = 1.0;
  else if (ay <= az)
    t[1] = 1.0;
  else
    t[2] = 1.0;
  double d = (t[0] * mu[0] + t[1] * mu[1]) + t[2] * mu[2];
  double u0 = t[0] - d * mu[0], u1 = t[1] - d * mu[1], u2 = t[2] - d * mu[2];
  double nrm = sqrt((u0 * u0 + u1 * u1) + u2 * u2);
  e1[0] = u0 / nrm;
  e1[1] = u1 / nrm;
  e1[2] = u2 / nrm;
  e2[0] = mu[1] * e1[2] - mu[2] * e1[1];
  e2[1] = mu[2] * e1[0] - mu[0] * e1[2];
  e2[2] = mu[0] * e1[1] - mu[1] * e1[0];
}

/* ===================================================================== */
/* Capsule distance                                                        */
/* ===================================================================== */
/* Closest approach of the two axis segments, centred parametrisation
 * P = r1 + s h1, Q = r2 + shift + t h2, s,t in [-1,1], h = (l/2) c.
 * Follows the role of distance_cylinder_capped (inc/intersect/
 * intersect_cylinder_capped.h:69-147) and cnt_intersec_shr
 * (inc/intersection_dev.h:10-80) but with the exact segment-segment
 * minimisation (clamp s, solve t, clamp t and re-solve s; Ericson,
 * Real-Time Collision Detection 5.1.9) instead of the reference's l/2+a
 * clamp slack and single-pass clamp.  Returns |P-Q|^2. */
static const double REACH_PAD = 1.000001;

double orc_pair_d2(const float* ri, const float* rj, const double shift[3], double R,
                   int* within) {
  double wx = ((double)ri[0] - (double)rj[0]) - shift[0];
  double wy = ((double)ri[1] - (double)rj[1]) - shift[1];
  double wz = ((double)ri[2] - (double)rj[2]) - shift[2];
  double hli = (double)ri[6] * 0.5, hlj = (double)rj[6] * 0.5;
  double rr = (hli + hlj) * REACH_PAD + R;
  double w2 = (wx * wx + wy * wy) + wz * wz;
  *within = (w2 <= rr * rr);
  double h1x = hli * (double)ri[3], h1y = hli * (double)ri[4], h1z = hli * (double)ri[5];
  double h2x = hlj * (double)rj[3], h2y = hlj * (double)rj[4], h2z = hlj * (double)rj[5];
  double a = (h1x * h1x + h1y * h1y) + h1z * h1z;
  double e = (h2x * h2x + h2y * h2y) + h2z * h2z;
  double b = (h1x * h2x + h1y * h2y) + h1z * h2z;
  double c = (h1x * wx + h1y * wy) + h1z * wz;
  double f = (h2x * wx + h2y * wy) + h2z * wz;
  double denom = a * e - b * b;
  double s = 0.0;
  if (denom > 0.0) {
    s = (b * f - c * e) / denom;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
  }
  double t = (b * s + f) / e;
  if (t < -1.0) {
    t = -1.0;
    s = (-b - c) / a;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
  } else if (t > 1.0) {
    t = 1.0;
    s = (b - c) / a;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
  }
  double vx = (wx + s * h1x) - t * h2x;
  double vy = (wy + s * h1y) - t * h2y;
  double vz = (wz + s * h1z) - t * h2z;
  return (vx * vx + vy * vy) + vz * vz;
}

/* Batch form for the tests: cut[i] = within && (contact ? d2 <= R^2 : d2 < R^2). */
void orc_pair_batch(int n, const float* ra, const float* rb, const double* sh, int contact,
                    double sep_or_gap, double soft_frac, double* d2out, int* within_out,
                    int* cut_out) {
  for (int i = 0; i < n; ++i) {
    const float* a = ra + 8 * (size_t)i;
    const float* b = rb + 8 * (size_t)i;
    double R = contact ? ((double)a[7] + (double)b[7]) + sep_or_gap
                       : soft_frac * ((double)a[7] + (double)b[7]) + sep_or_gap;
    int w;
    double d2 = orc_pair_d2(a, b, &sh[3 * (size_t)i], R, &w);
    d2out[i] = d2;
    within_out[i] = w;
    cut_out[i] = w && (contact ? (d2 <= R * R) : (d2 < R * R));
  }
}

/* Hard-core distance: soft_frac*(a1+a2)+sep (inc/intersection_dev.h:78,
 * inc/intersect/intersect_cylinder_capped.h:157-159 without the +1e-6). */
static double hard_core(const nf_params* p, const float* ri, const float* rj) {
  return p->soft_frac * ((double)ri[7] + (double)rj[7]) + p->sep;
}

/* Contact (tunnelling) area S of the pair (ri, rj + shift): the area model
 * of surf_to_cylinder_capped (inc/intersect/intersect_cylinder_capped.h:
 * 204-293) on the line geometry of distance_cylinder_capped (:69-147), as
 * get_surface_cyl2cyl applies it per contact (src/get_contact_data.cu:
 * 130-172), in FP64 with unit directions and t in length units:
 *  - t1, t2: closest points of the infinite axes (:82-105; within 1e-6 of
 *    parallel, t1 = +-l1/2 towards rod j and t2 its projection), D their
 *    distance (Dtt.x);
 *  - point disc (:224-233): h1 = max(D - a2 - rlim, 0), h2 = max(D - a1 -
 *    rlim, 0), dr_i = a_i sqrt(1 - (h_i/a_i)^2) when h_i <= a_i, dr =
 *    min(dr1, dr2), S = pi dr^2;
 *  - crossing ellipse (:237-266) when |t1| < l1/2 - a1 and |t2| < l2/2 - a2:
 *    A+- = max(dr, min(dr2 / sin(gamma/2), rod-end distances along the
 *    bisector ax1 of c1 and c2' = c2 sign(c1.c2))), S = max(S, pi/2 dr (A+ +
 *    A-)); the end distances are measured from the contact points, and rod
 *    j's parameter along c2' is t2 sign(c1.c2) (the reference adds the rod
 *    centres to the end distances, :257-260, and uses copysign(t2, c1.c2),
 *    :253);
 *  - near-parallel strip (:268-290) when sin gamma <= max(a)/min(l) and
 *    |t1| <= l1/2 + a1 or |t2| <= l2/2 + a2: the rod ends projected on c1
 *    (relative to rod i's centre), S = max(S, |p2 - p1| 2 dr). */
static double dmn(double a, double b) { return b < a ? b : a; }
static double dmx(double a, double b) { return b > a ? b : a; }

double orc_contact_area(const float* ri, const float* rj, const double shift[3], double rlim) {
  const double PI = 3.14159265358979323846;
  const double l1 = (double)ri[6], l2 = (double)rj[6];
  const double a1 = (double)ri[7], a2 = (double)rj[7];
  const double c1[3] = {(double)ri[3], (double)ri[4], (double)ri[5]};
  const double c2[3] = {(double)rj[3], (double)rj[4], (double)rj[5]};
  double b[3];
  for (int k = 0; k < 3; ++k) b[k] = ((double)rj[k] - (double)ri[k]) + shift[k];
  const double d = (c1[0] * c2[0] + c1[1] * c2[1]) + c1[2] * c2[2];
  const double bc1 = (b[0] * c1[0] + b[1] * c1[1]) + b[2] * c1[2];
  const double bc2 = (b[0] * c2[0] + b[1] * c2[1]) + b[2] * c2[2];
  double t1, t2;
  if (1.0 - fabs(d) > 1e-6) {
    const double den = d * d - 1.0;
    t1 = (d * bc2 - bc1) / den;
    t2 = (bc2 - d * bc1) / den;
  } else {
    t1 = bc1 < 0.0 ? -(l1 * 0.5) : l1 * 0.5;
    t2 = d * t1 - bc2;
  }
  double v2 = 0.0;
  double v[3];
  for (int k = 0; k < 3; ++k) v[k] = (t1 * c1[k] - t2 * c2[k]) - b[k];
  v2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
  const double D = sqrt(v2);
  const double h1 = dmx((D - a2) - rlim, 0.0), h2 = dmx((D - a1) - rlim, 0.0);
  double dr1 = 0.0, dr2 = 0.0;
  if (h1 <= a1) {
    const double q = h1 / a1;
    dr1 = a1 * sqrt(dmx(1.0 - q * q, 0.0));
  }
  if (h2 <= a2) {
    const double q = h2 / a2;
    dr2 = a2 * sqrt(dmx(1.0 - q * q, 0.0));
  }
  const double dr = dmn(dr1, dr2);
  double S = (PI * dr) * dr;
  const double sg = d < 0.0 ? -1.0 : 1.0;
  const double dp = d * sg;
  const double sng = sqrt(dmx(1.0 - dp * dp, 0.0));
  const double at1 = fabs(t1), at2 = fabs(t2);
  if (at1 < l1 * 0.5 - a1 && at2 < l2 * 0.5 - a2) {
    double ax[3] = {c1[0] + sg * c2[0], c1[1] + sg * c2[1], c1[2] + sg * c2[2]};
    const double an = sqrt((ax[0] * ax[0] + ax[1] * ax[1]) + ax[2] * ax[2]);
    for (int k = 0; k < 3; ++k) ax[k] = ax[k] / an;
    const double csg = sqrt(dmx(1.0 - sng * sng, 0.0));
    const double sngh = sqrt((1.0 - csg) * 0.5);
    const double t2p = sg * t2;
    const double k1 = (c1[0] * ax[0] + c1[1] * ax[1]) + c1[2] * ax[2];
    const double k2 = ((sg * c2[0]) * ax[0] + (sg * c2[1]) * ax[1]) + (sg * c2[2]) * ax[2];
    const double dt0 = sngh > 1e-6 ? dr2 / sngh : l1 * 0.5;
    const double dt1p = (l1 * 0.5 - t1) * k1;
    const double dt1m = fabs((-(l1 * 0.5) - t1) * k1);
    const double dt2p = (l2 * 0.5 - t2p) * k2;
    const double dt2m = fabs((-(l2 * 0.5) - t2p) * k2);
    const double Ap = dmx(dr, dmn(dt0, dmn(dt1p, dt2p)));
    const double Am = dmx(dr, dmn(dt0, dmn(dt1m, dt2m)));
    S = dmx(S, ((0.5 * PI) * dr) * (Ap + Am));
  }
  if (sng <= dmx(a1, a2) / dmn(l1, l2) && (at1 <= l1 * 0.5 + a1 || at2 <= l2 * 0.5 + a2)) {
    const double t1a = bc1 < 0.0 ? -(l1 * 0.5) : l1 * 0.5;
    const double t2a = bc2 < 0.0 ? -(l2 * 0.5) : l2 * 0.5;
    const double p1a = t1a, p1b = -t1a;
    const double p2a = bc1 + t2a * d, p2b = bc1 - t2a * d;
    const double p1 = bc1 > 0.0 ? dmn(p1a, p2a) : dmx(p1a, p2a);
    const double p2 = bc1 > 0.0 ? dmx(p1b, p2b) : dmn(p1b, p2b);
    S = dmx(S, (fabs(p2 - p1) * 2.0) * dr);
  }
  return S;
}

/* Contact areas of a contact list (image packs J's shift as signed bytes). */
void orc_contact_areas(const nf_params* p, const float* rods, const uint32_t* I, const uint32_t* J,
                       const int32_t* image, int64_t nc, double rlim, double* out) {
  for (int64_t k = 0; k < nc; ++k) {
    double sh[3];
    for (int a = 0; a < 3; ++a)
      sh[a] = (double)(int)(int8_t)((image[k] >> (8 * a)) & 0xFF) * p->box[a];
    out[k] = orc_contact_area(&rods[8 * (size_t)I[k]], &rods[8 * (size_t)J[k]], sh, rlim);
  }
}

/* ===================================================================== */
/* Candidate sampling                                                      */
/* ===================================================================== */
/* Von Neumann acceptance against the density target and orientation from
 * the target law, the job of cuda_CNTgenerate_orient (inc/cuda_CNTgenerate.h:
 * 136-260: uniform position, accept if u < density(x) at :227-232,
 * genCosine/inc_rotation at :262-295).  Isotropic directions use Marsaglia's
 * (1972) disk method; vMF uses Wood's (1994) exact inverse CDF for p = 3;
 * log-normal lengths use Marsaglia's polar method. */
static int self_image_overlap(const nf_params* p, const float* r) {
  double R = hard_core(p, r, r);
  double rr = ((double)r[6] * 0.5 + (double)r[6] * 0.5) * REACH_PAD + R;
  double lmin = p->box[0];
  if (p->box[1] < lmin) lmin = p->box[1];
  if (p->box[2] < lmin) lmin = p->box[2];
  if (rr < lmin) return 0;
  int K[3];
  for (int d = 0; d < 3; ++d) K[d] = (int)floor(rr / p->box[d]) + 1;
  for (int nz = -K[2]; nz <= K[2]; ++nz)
    for (int ny = -K[1]; ny <= K[1]; ++ny)
      for (int nx = -K[0]; nx <= K[0]; ++nx) {
        if (nx == 0 && ny == 0 && nz == 0) continue;
        double sh[3] = {(double)nx * p->box[0], (double)ny * p->box[1], (double)nz * p->box[2]};
        int within;
        double d2 = orc_pair_d2(r, r, sh, R, &within);
        if (within && d2 < R * R) return 1;
      }
  return 0;
}

/* Density map value: trilinear interpolation with periodic wrap, mesh points
 * at the centres of the map's cells, as the reference's normalised, linearly
 * filtered, wrapping texture (src/den2tex.h:5-53; read at
 * inc/cuda_CNTgenerate.h:228), in FP64; interpolation along z, then y,
 * then x. */
double orc_mesh_density(const nf_params* p, double x, double y, double z) {
  const float* v = p->den_mesh;
  const int32_t* e = p->den_ext;
  double pos[3] = {x, y, z};
  int i0[3], i1[3];
  double f[3];
  for (int a = 0; a < 3; ++a) {
    double t = (pos[a] / p->box[a]) * (double)e[a] - 0.5;
    double fl = floor(t);
    f[a] = t - fl;
    int k = (int)fl;
    if (k < 0) k += e[a];
    if (k >= e[a]) k -= e[a];
    i0[a] = k;
    i1[a] = k + 1 == e[a] ? 0 : k + 1;
  }
  uint64_t sy = (uint64_t)e[2], sx = (uint64_t)e[1] * (uint64_t)e[2];
  uint64_t x0 = (uint64_t)i0[0] * sx, x1 = (uint64_t)i1[0] * sx;
  uint64_t y0 = (uint64_t)i0[1] * sy, y1 = (uint64_t)i1[1] * sy;
  uint64_t z0 = (uint64_t)i0[2], z1 = (uint64_t)i1[2];
  double gz = 1.0 - f[2], gy = 1.0 - f[1], gx = 1.0 - f[0];
  double c00 = (double)v[x0 + y0 + z0] * gz + (double)v[x0 + y0 + z1] * f[2];
  double c01 = (double)v[x0 + y1 + z0] * gz + (double)v[x0 + y1 + z1] * f[2];
  double c10 = (double)v[x1 + y0 + z0] * gz + (double)v[x1 + y0 + z1] * f[2];
  double c11 = (double)v[x1 + y1 + z0] * gz + (double)v[x1 + y1 + z1] * f[2];
  double c0 = c00 * gy + c01 * f[1];
  double c1 = c10 * gy + c11 * f[1];
  return c0 * gx + c1 * f[0];
}

int orc_candidate(const nf_params* p, const double* agg, uint32_t round, uint32_t slot,
                  float out[8]) {
  uint32_t o[4];
  double Lx = p->box[0], Ly = p->box[1], Lz = p->box[2];
  blk(p, slot, round, 0u, 0u, o);
  double ux = orc_u01(o[0], o[1]), uy = orc_u01(o[2], o[3]);
  blk(p, slot, round, 1u, 0u, o);
  double uz = orc_u01(o[0], o[1]), ud = orc_u01(o[2], o[3]);
  float xf = (float)(ux * Lx), yf = (float)(uy * Ly), zf = (float)(uz * Lz);
  if (!(xf < (float)Lx)) xf = 0.0f;
  if (!(yf < (float)Ly)) yf = 0.0f;
  if (!(zf < (float)Lz)) zf = 0.0f;

  if (p->den_law == NF_DEN_AGGLOM) {
    double x = (double)xf, y = (double)yf, z = (double)zf;
    double d2min = INFINITY;
    for (int k = 0; k < p->n_agg; ++k) {
      double dx = x - agg[3 * k + 0], dy = y - agg[3 * k + 1], dz = z - agg[3 * k + 2];
      if (dx > 0.5 * Lx) dx -= Lx; else if (dx < -0.5 * Lx) dx += Lx;
      if (dy > 0.5 * Ly) dy -= Ly; else if (dy < -0.5 * Ly) dy += Ly;
      if (dz > 0.5 * Lz) dz -= Lz; else if (dz < -0.5 * Lz) dz += Lz;
      double d2 = (dx * dx + dy * dy) + dz * dz;
      if (d2 < d2min) d2min = d2;
    }
    double inv2s2 = 1.0 / (2.0 * p->agg_sigma * p->agg_sigma);
    double rho = 1.0 + (p->agg_peak - 1.0) * orc_exp(-(d2min * inv2s2));
    if (!(ud * p->agg_peak < rho)) return 0;
  } else if (p->den_law == NF_DEN_MESH) {
    double rho = orc_mesh_density(p, (double)xf, (double)yf, (double)zf);
    if (!(ud < rho)) return 0;
  }

  double cx = 0.0, cy = 0.0, cz = 0.0;
  int ok = 0;
  if (p->ort_law == NF_ORT_VMF) {
    double e1[3], e2[3];
    orc_vmf_basis(p->axis, e1, e2);
    blk(p, slot, round, 8u, 0u, o);
    double uw = orc_u01(o[0], o[1]);
    double em2k = orc_exp(-2.0 * p->kappa);
    double w = 1.0 + orc_log(uw + (1.0 - uw) * em2k) / p->kappa;
    if (w > 1.0) w = 1.0;
    if (w < -1.0) w = -1.0;
    for (uint32_t t = 0; t < 32 && !ok; ++t) {
      blk(p, slot, round, 9u + t, 0u, o);
      double v1 = 2.0 * orc_u01(o[0], o[1]) - 1.0, v2 = 2.0 * orc_u01(o[2], o[3]) - 1.0;
      double s = v1 * v1 + v2 * v2;
      if (s < 1.0 && s > 0.0) {
        double rr = sqrt(1.0 - w * w) / sqrt(s);
        double a1 = v1 * rr, a2 = v2 * rr;
        cx = (w * p->axis[0] + a1 * e1[0]) + a2 * e2[0];
        cy = (w * p->axis[1] + a1 * e1[1]) + a2 * e2[1];
        cz = (w * p->axis[2] + a1 * e1[2]) + a2 * e2[2];
        ok = 1;
      }
    }
  } else {
    for (uint32_t t = 0; t < 32 && !ok; ++t) {
      blk(p, slot, round, 8u + t, 0u, o);
      double v1 = 2.0 * orc_u01(o[0], o[1]) - 1.0, v2 = 2.0 * orc_u01(o[2], o[3]) - 1.0;
      double s = v1 * v1 + v2 * v2;
      if (s < 1.0 && s > 0.0) {
        double q = 2.0 * sqrt(1.0 - s);
        cx = v1 * q;
        cy = v2 * q;
        cz = 1.0 - 2.0 * s;
        ok = 1;
      }
    }
  }
  if (!ok) return -1;

  double L = p->len;
  if (p->len_law == NF_LEN_LOGNORMAL) {
    ok = 0;
    double z = 0.0;
    for (uint32_t t = 0; t < 32 && !ok; ++t) {
      blk(p, slot, round, 48u + t, 0u, o);
      double v1 = 2.0 * orc_u01(o[0], o[1]) - 1.0, v2 = 2.0 * orc_u01(o[2], o[3]) - 1.0;
      double s = v1 * v1 + v2 * v2;
      if (s < 1.0 && s > 0.0) {
        z = v1 * sqrt(-2.0 * orc_log(s) / s);
        ok = 1;
      }
    }
    if (!ok) return -1;
    L = p->len * orc_exp(p->len_sigma * z);
    if (L < p->len_min) L = p->len_min;
    if (L > p->len_max) L = p->len_max;
  }
  out[0] = xf;
  out[1] = yf;
  out[2] = zf;
  out[3] = (float)cx;
  out[4] = (float)cy;
  out[5] = (float)cz;
  out[6] = (float)L;
  out[7] = (float)p->radius;
  if (self_image_overlap(p, out)) return -2;
  return 1;
}

/* ===================================================================== */
/* Periodic cell grid (oracle's own neighbour search)                      */
/* ===================================================================== */
typedef struct {
  int n[3];
  double L[3];
  int ncells;
  int nst;
  int* st; /* stencil offsets, 3 ints each */
} grid_t;

static void grid_init(grid_t* g, const double L[3], double reach) {
  double target = reach * 0.5;
  for (int d = 0; d < 3; ++d) {
    g->L[d] = L[d];
    double nd = floor(L[d] / target);
    if (nd < 1.0) nd = 1.0;
    if (nd > 160.0) nd = 160.0;
    g->n[d] = (int)nd;
  }
  g->ncells = g->n[0] * g->n[1] * g->n[2];
  double rm = reach + 1e-5;
  int k[3];
  for (int d = 0; d < 3; ++d) k[d] = (int)floor(rm / (L[d] / g->n[d])) + 1;
  g->st = (int*)malloc(sizeof(int) * 3 * (2 * k[0] + 1) * (2 * k[1] + 1) * (2 * k[2] + 1));
  g->nst = 0;
  for (int dz = -k[2]; dz <= k[2]; ++dz)
    for (int dy = -k[1]; dy <= k[1]; ++dy)
      for (int dx = -k[0]; dx <= k[0]; ++dx) {
        int dd[3] = {dx, dy, dz};
        double m2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          int a = abs(dd[d]) - 1;
          if (a > 0) {
            double t = a * (L[d] / g->n[d]);
            m2 += t * t;
          }
        }
        if (m2 <= rm * rm) {
          g->st[3 * g->nst + 0] = dx;
          g->st[3 * g->nst + 1] = dy;
          g->st[3 * g->nst + 2] = dz;
          g->nst++;
        }
      }
}

static void grid_free(grid_t* g) { free(g->st); }

static void grid_coord(const grid_t* g, const float* r, int c[3]) {
  for (int d = 0; d < 3; ++d) {
    int i = (int)((double)r[d] * (double)g->n[d] / g->L[d]);
    if (i < 0) i = 0;
    if (i >= g->n[d]) i = g->n[d] - 1;
    c[d] = i;
  }
}

static int grid_cell(const grid_t* g, const float* r) {
  int c[3];
  grid_coord(g, r, c);
  return c[0] + g->n[0] * (c[1] + g->n[1] * c[2]);
}

/* neighbour cell of c at stencil offset s: returns cell id, image in img */
static int grid_neigh(const grid_t* g, const int c[3], const int* s, int img[3]) {
  int m[3];
  for (int d = 0; d < 3; ++d) {
    int raw = c[d] + s[d];
    int mm = raw % g->n[d];
    if (mm < 0) mm += g->n[d];
    img[d] = (raw - mm) / g->n[d];
    m[d] = mm;
  }
  return m[0] + g->n[0] * (m[1] + g->n[1] * m[2]);
}

/* ===================================================================== */
/* Generation rounds                                                       */
/* ===================================================================== */
/* Round schedule: the number of candidate slots per round is an integer
 * function of the round history, identical on CPU and GPU.  The reference
 * sizes rounds from per-cell quotas instead (src/cudaMain.cu:219-351,
 * src/simPrms.cpp:881-911). */
static uint32_t clamp_batch(uint64_t m, const nf_params* p) {
  if (m < p->min_batch) m = p->min_batch;
  if (m > p->max_batch) m = p->max_batch;
  return (uint32_t)m;
}
static uint32_t sched_first(const nf_params* p, uint32_t deficit) {
  return clamp_batch((uint64_t)deficit + (uint64_t)(deficit / 4u), p);
}
static uint32_t sched_next(const nf_params* p, uint32_t deficit, uint32_t mprev, uint32_t aprev) {
  if (aprev == 0) return p->max_batch;
  uint64_t num = (uint64_t)deficit * (uint64_t)mprev * 5ull;
  uint64_t den = (uint64_t)aprev * 4ull;
  return clamp_batch((num + den - 1) / den, p);
}

typedef struct {
  float r[8];
  uint32_t slot;
  int cell;
  int dead1, dead2;
} prop_t;

/* Round rule (replaces the launch-order dependent removal of
 * cuda_InternalIsect_noLck / cuda_NewExtIsectLB_parted / cuda_OldExtIsectLB,
 * inc/intersection_kern.h:12-364): a proposal is rejected if it overlaps an
 * accepted rod (stage 1), or overlaps a stage-1 survivor with a smaller slot
 * (stage 2).  Survivors are accepted in slot order up to the target; ids
 * continue the accepted count.  Termination follows the loop condition of
 * cudaMain (src/cudaMain.cu:113-114): target reached, round cap, no
 * progress (here: stall_rounds rounds without an acceptance). */
int orc_generate(const nf_params* p, float* rods, orc_gen_stats* st, uint32_t* round_accepts,
                 int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  memset(st, 0, sizeof(*st));
  uint32_t N = p->n_rods;
  if (N == 0 || p->max_batch == 0 || p->min_batch == 0) return NF_E_INVALID;
  double* agg = NULL;
  if (p->den_law == NF_DEN_AGGLOM) {
    agg = (double*)malloc(sizeof(double) * 3 * (p->n_agg > 0 ? p->n_agg : 1));
    orc_agg_centres(p, agg);
  }
  double lmax = p->len_law == NF_LEN_LOGNORMAL ? (double)(float)p->len_max : (double)(float)p->len;
  double amax = (double)(float)p->radius;
  double Rmax = p->soft_frac * (amax + amax) + p->sep;
  double reach = (lmax * 0.5 + lmax * 0.5) * REACH_PAD + Rmax;
  grid_t g;
  grid_init(&g, p->box, reach);
  int* head = (int*)malloc(sizeof(int) * g.ncells);
  int* next = (int*)malloc(sizeof(int) * N);
  int* phead = (int*)malloc(sizeof(int) * g.ncells);
  for (int i = 0; i < g.ncells; ++i) head[i] = -1;

  uint32_t n_acc = 0, round = 0, stall = 0, mprev = 0, aprev = 0;
  int status = NF_STATUS_RUNNING;
  prop_t* props = NULL;
  size_t props_cap = 0;
  int* pnext = NULL;
  while (status == NF_STATUS_RUNNING) {
    uint32_t deficit = N - n_acc;
    uint32_t M = round == 0 ? sched_first(p, deficit) : sched_next(p, deficit, mprev, aprev);
    if (props_cap < M) {
      props_cap = M;
      props = (prop_t*)realloc(props, sizeof(prop_t) * props_cap);
      pnext = (int*)realloc(pnext, sizeof(int) * props_cap);
    }
    /* candidates */
    int P = 0;
    for (uint32_t s = 0; s < M; ++s) {
      float r[8];
      int v = orc_candidate(p, agg, round, s, r);
      if (v == 1) {
        memcpy(props[P].r, r, sizeof(r));
        props[P].slot = s;
        props[P].cell = grid_cell(&g, r);
        props[P].dead1 = props[P].dead2 = 0;
        ++P;
      }
    }
    uint64_t tests = 0, rej_old = 0, rej_new = 0;
    /* stage 1: proposals vs accepted rods */
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : tests, rej_old)
    for (int k = 0; k < P; ++k) {
      prop_t* q = &props[k];
      int c[3], img[3];
      grid_coord(&g, q->r, c);
      int dead = 0;
      for (int si = 0; si < g.nst; ++si) {
        int nc = grid_neigh(&g, c, &g.st[3 * si], img);
        double sh[3] = {(double)img[0] * g.L[0], (double)img[1] * g.L[1], (double)img[2] * g.L[2]};
        for (int j = head[nc]; j >= 0; j = next[j]) {
          const float* rj = &rods[8 * (size_t)j];
          double R = hard_core(p, q->r, rj);
          int within;
          double d2 = orc_pair_d2(q->r, rj, sh, R, &within);
          if (within) {
            ++tests;
            if (d2 < R * R) dead = 1;
          }
        }
      }
      q->dead1 = dead;
      rej_old += (uint64_t)dead;
    }
    /* stage 2: survivors vs higher-priority survivors */
    for (int i = 0; i < g.ncells; ++i) phead[i] = -1;
    for (int k = 0; k < P; ++k)
      if (!props[k].dead1) {
        pnext[k] = phead[props[k].cell];
        phead[props[k].cell] = k;
      }
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : tests, rej_new)
    for (int k = 0; k < P; ++k) {
      prop_t* q = &props[k];
      if (q->dead1) continue;
      int c[3], img[3];
      grid_coord(&g, q->r, c);
      int dead = 0;
      for (int si = 0; si < g.nst; ++si) {
        int nc = grid_neigh(&g, c, &g.st[3 * si], img);
        double sh[3] = {(double)img[0] * g.L[0], (double)img[1] * g.L[1], (double)img[2] * g.L[2]};
        for (int j = phead[nc]; j >= 0; j = pnext[j]) {
          if (props[j].slot >= q->slot) continue;
          double R = hard_core(p, q->r, props[j].r);
          int within;
          double d2 = orc_pair_d2(q->r, props[j].r, sh, R, &within);
          if (within) {
            ++tests;
            if (d2 < R * R) dead = 1;
          }
        }
      }
      q->dead2 = dead;
      rej_new += (uint64_t)dead;
    }
    /* commit in slot order */
    uint32_t acc = 0;
    for (int k = 0; k < P && acc < deficit; ++k) {
      if (props[k].dead1 || props[k].dead2) continue;
      uint32_t id = n_acc + acc;
      memcpy(&rods[8 * (size_t)id], props[k].r, sizeof(float) * 8);
      int cell = props[k].cell;
      next[id] = head[cell];
      head[cell] = (int)id;
      ++acc;
    }
    st->candidates += M;
    st->proposals += (uint64_t)P;
    st->pair_tests += tests;
    st->rejected_old += rej_old;
    st->rejected_new += rej_new;
    if (round_accepts && round < p->max_rounds) round_accepts[round] = acc;
    n_acc += acc;
    mprev = M;
    aprev = acc;
    round += 1;
    stall = acc == 0 ? stall + 1 : 0;
    if (n_acc == N)
      status = NF_STATUS_TARGET;
    else if (stall >= p->stall_rounds)
      status = NF_STATUS_STALLED;
    else if (round >= p->max_rounds)
      status = NF_STATUS_MAX_ROUNDS;
  }
  st->n_accepted = n_acc;
  st->rounds = round;
  st->status = status;
  free(props);
  free(pnext);
  free(head);
  free(next);
  free(phead);
  free(agg);
  grid_free(&g);
  return NF_OK;
}

/* ===================================================================== */
/* Contacts                                                                */
/* ===================================================================== */
typedef struct {
  uint32_t I, J;
  int32_t img;
  float g32;
  double g64;
} cnt_t;

static int32_t pack_img(const int img[3]) {
  return (int32_t)(((uint32_t)(img[0] & 0xFF)) | ((uint32_t)(img[1] & 0xFF) << 8) |
                   ((uint32_t)(img[2] & 0xFF) << 16));
}
static void unpack_img(int32_t c, int img[3]) {
  img[0] = (int)(int8_t)(c & 0xFF);
  img[1] = (int)(int8_t)((c >> 8) & 0xFF);
  img[2] = (int)(int8_t)((c >> 16) & 0xFF);
}
static int img_cmp(int32_t a, int32_t b) {
  int x[3], y[3];
  unpack_img(a, x);
  unpack_img(b, y);
  for (int d = 0; d < 3; ++d)
    if (x[d] != y[d]) return x[d] < y[d] ? -1 : 1;
  return 0;
}
static int cnt_cmp(const void* pa, const void* pb) {
  const cnt_t* a = (const cnt_t*)pa;
  const cnt_t* b = (const cnt_t*)pb;
  if (a->I != b->I) return a->I < b->I ? -1 : 1;
  if (a->J != b->J) return a->J < b->J ? -1 : 1;
  return img_cmp(a->img, b->img);
}

/* Contact detection: the CONTACT_COUNT / CONTACT_RECORD modes of
 * core_cell_interact_kernel (src/core_cell_interaction_kernel.cu:70-176,
 * 181-315) with the upper-triangular rule (neigID >= selfID, J > i; :115,
 * :246-251) expressed on rod ids, contact test d < a1+a2+gap
 * (inc/intersect/intersect_cylinder_capped.h:157-159, separation =
 * STAT_DIST_LIM per core_cell_interaction_kernel.h:174-190), and the stored
 * distance being the surface gap d - a1 - a2 (the reference stores the axis
 * point distance, src/get_contact_data.cu:28-68). */
int64_t orc_contacts(const nf_params* p, const float* rods, uint32_t n, double gap,
                     uint32_t* I, uint32_t* J, float* gap32, double* gap64, int32_t* image,
                     int64_t cap, uint64_t* pair_tests, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  double lmax = 0.0, amax = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    if ((double)rods[8 * (size_t)i + 6] > lmax) lmax = rods[8 * (size_t)i + 6];
    if ((double)rods[8 * (size_t)i + 7] > amax) amax = rods[8 * (size_t)i + 7];
  }
  double reach = (lmax * 0.5 + lmax * 0.5) * REACH_PAD + ((amax + amax) + gap);
  grid_t g;
  grid_init(&g, p->box, reach);
  int* head = (int*)malloc(sizeof(int) * g.ncells);
  int* next = (int*)malloc(sizeof(int) * (n ? n : 1));
  for (int i = 0; i < g.ncells; ++i) head[i] = -1;
  for (int64_t i = (int64_t)n - 1; i >= 0; --i) {
    int cc = grid_cell(&g, &rods[8 * (size_t)i]);
    next[i] = head[cc];
    head[cc] = (int)i;
  }
  size_t cap_all = 1024, cnt = 0;
  cnt_t* all = (cnt_t*)malloc(sizeof(cnt_t) * cap_all);
  uint64_t tests = 0;
#pragma omp parallel
  {
    size_t lcap = 256, lcnt = 0;
    cnt_t* loc = (cnt_t*)malloc(sizeof(cnt_t) * lcap);
    uint64_t ltests = 0;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      const float* ri = &rods[8 * (size_t)i];
      int c[3], img[3];
      grid_coord(&g, ri, c);
      for (int si = 0; si < g.nst; ++si) {
        int nc = grid_neigh(&g, c, &g.st[3 * si], img);
        double sh[3] = {(double)img[0] * g.L[0], (double)img[1] * g.L[1], (double)img[2] * g.L[2]};
        for (int j = head[nc]; j >= 0; j = next[j]) {
          if ((int64_t)j <= i) continue;
          const float* rj = &rods[8 * (size_t)j];
          double R = ((double)ri[7] + (double)rj[7]) + gap;
          int within;
          double d2 = orc_pair_d2(ri, rj, sh, R, &within);
          if (!within) continue;
          ++ltests;
          if (d2 <= R * R) {
            if (lcnt == lcap) {
              lcap *= 2;
              loc = (cnt_t*)realloc(loc, sizeof(cnt_t) * lcap);
            }
            cnt_t* e = &loc[lcnt++];
            e->I = (uint32_t)i;
            e->J = (uint32_t)j;
            e->img = pack_img(img);
            e->g64 = sqrt(d2) - ((double)ri[7] + (double)rj[7]);
            e->g32 = (float)e->g64;
          }
        }
      }
    }
#pragma omp critical
    {
      if (cnt + lcnt > cap_all) {
        while (cnt + lcnt > cap_all) cap_all *= 2;
        all = (cnt_t*)realloc(all, sizeof(cnt_t) * cap_all);
      }
      memcpy(all + cnt, loc, sizeof(cnt_t) * lcnt);
      cnt += lcnt;
      tests += ltests;
    }
    free(loc);
  }
  qsort(all, cnt, sizeof(cnt_t), cnt_cmp);
  for (size_t k = 0; k < cnt && (int64_t)k < cap; ++k) {
    if (I) I[k] = all[k].I;
    if (J) J[k] = all[k].J;
    if (gap32) gap32[k] = all[k].g32;
    if (gap64) gap64[k] = all[k].g64;
    if (image) image[k] = all[k].img;
  }
  if (pair_tests) *pair_tests = tests;
  free(all);
  free(head);
  free(next);
  grid_free(&g);
  return (int64_t)cnt;
}

/* ===================================================================== */
/* Histograms                                                              */
/* ===================================================================== */
/* Angle bins by comparison with cos(bin edge) so no inverse cosine is
 * needed: bin = #{k in 1..nb-1 : cos(theta) <= cos(k*range/nb)}.  The
 * reference bins acosf(...) with atomicInc-wrapped uint32 counters
 * (inc/statKernel_ANGL.h:58-66, inc/statKernel_MUT.h:99-115); counts here
 * are exact int64. */
static int angle_bin(double cosv, const double* edges, int nb) {
  int b = 0;
  for (int k = 1; k < nb; ++k)
    if (cosv <= edges[k]) b++;
  return b;
}

int orc_histograms(const nf_params* p, const float* rods, uint32_t n, const uint32_t* I,
                   const uint32_t* J, int64_t nc, const nf_hist_spec* spec, int64_t* centre,
                   int64_t* polar, int64_t* mutual, int64_t* phitheta) {
  int nb = spec->bins;
  if (nb <= 0 || spec->centre_axis < 0 || spec->centre_axis > 2) return NF_E_INVALID;
  double PI = 3.14159265358979323846;
  double* ep = (double*)malloc(sizeof(double) * (nb + 1));
  double* em = (double*)malloc(sizeof(double) * (nb + 1));
  for (int k = 0; k <= nb; ++k) {
    ep[k] = cos((double)k * PI / (double)nb);
    em[k] = cos((double)k * (0.5 * PI) / (double)nb);
  }
  double mu[3] = {spec->pol_axis[0], spec->pol_axis[1], spec->pol_axis[2]};
  double mn = sqrt((mu[0] * mu[0] + mu[1] * mu[1]) + mu[2] * mu[2]);
  mu[0] = mu[0] / mn;
  mu[1] = mu[1] / mn;
  mu[2] = mu[2] / mn;
  int ax = spec->centre_axis;
  double L = p->box[ax];
  for (int k = 0; k < nb; ++k) centre[k] = polar[k] = 0;
  if (mutual)
    for (int k = 0; k < nb; ++k) mutual[k] = 0;
  int npb = spec->phi_bins, ntb = spec->theta_bins;
  double* eth = NULL;
  if (phitheta && npb > 0 && ntb > 0) {
    for (int k = 0; k < npb * ntb; ++k) phitheta[k] = 0;
    eth = (double*)malloc(sizeof(double) * (ntb + 1));
    for (int k = 0; k <= ntb; ++k) eth[k] = cos((double)k * PI / (double)ntb);
  }
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = &rods[8 * (size_t)i];
    double t = (double)r[ax] / L;
    int b = (int)(t * (double)nb);
    if (b < 0) b = 0;
    if (b >= nb) b = nb - 1;
    centre[b]++;
    double cx = r[3], cy = r[4], cz = r[5];
    double cn = sqrt((cx * cx + cy * cy) + cz * cz);
    double cosv = ((cx * mu[0] + cy * mu[1]) + cz * mu[2]) / cn;
    polar[angle_bin(cosv, ep, nb)]++;
    if (eth) {
      /* phi-theta as statKernel_ANGL: theta from +z, phi from +x of the
       * projection over [0, 2pi) (the reference folds phi to [0, pi],
       * inc/statKernel_ANGL.h:60-65); phi bin by the same edge rule on the
       * half-plane: bin in [0, npb/2) for y >= 0, mirrored otherwise. */
      double ct = cz / cn;
      int tb = angle_bin(ct, eth, ntb);
      double rxy = sqrt(cx * cx + cy * cy);
      int pb = 0;
      if (rxy > 0.0) {
        double cp = cx / rxy;
        int half = npb / 2;
        double PIh = PI;
        int bb = 0;
        for (int k = 1; k < half; ++k)
          if (cp <= cos((double)k * PIh / (double)half)) bb++;
        pb = cy >= 0.0 ? bb : (npb - 1 - bb);
      }
      phitheta[(size_t)tb * npb + pb]++;
    }
  }
  if (mutual && I && J) {
    for (int64_t k = 0; k < nc; ++k) {
      const float* a = &rods[8 * (size_t)I[k]];
      const float* b = &rods[8 * (size_t)J[k]];
      double ax_ = a[3], ay_ = a[4], az_ = a[5], bx = b[3], by = b[4], bz = b[5];
      double na = sqrt((ax_ * ax_ + ay_ * ay_) + az_ * az_);
      double nbv = sqrt((bx * bx + by * by) + bz * bz);
      double d = fabs((ax_ * bx + ay_ * by) + az_ * bz) / (na * nbv);
      mutual[angle_bin(d, em, nb)]++;
    }
  }
  free(ep);
  free(em);
  free(eth);
  return NF_OK;
}

/* ===================================================================== */
/* Network                                                                 */
/* ===================================================================== */
static uint32_t uf_find(uint32_t* par, uint32_t x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

typedef struct {
  uint32_t col;
  int32_t img;
  double g;
} ent_t;
static int ent_cmp(const void* pa, const void* pb) {
  const ent_t* a = (const ent_t*)pa;
  const ent_t* b = (const ent_t*)pb;
  if (a->col != b->col) return a->col < b->col ? -1 : 1;
  return img_cmp(a->img, b->img);
}

/* Blocked sequential sum: 1024-entry blocks summed in order, then the
 * block sums in order. */
static double blocked_sum(const double* v, size_t n) {
  double tot = 0.0;
  for (size_t b = 0; b < n; b += 1024) {
    double s = 0.0;
    size_t e = b + 1024 < n ? b + 1024 : n;
    for (size_t k = b; k < e; ++k) s += v[k];
    tot += s;
  }
  return tot;
}

/* Dead-end pruning: the `reduce` verb (src/NanoFiller.cpp:71-90) driving
 * reduce() (src/conductivity.cu:79-138): each iteration recounts every
 * present rod's contacts among the present rods (contact_count_kernel_full,
 * inc/conductKernels.h:40-113) and removes, all at once, the rods whose
 * count is below min_cont (reductKernel, :116-162); iterations repeat until
 * one removes nothing (counted, as the reference's iter).  Contacts are
 * the list entries whose image does not cross the electrode axis; a rod's
 * coupling to an electrode plane (gap to the plane <= gcut, the rule of
 * orc_network) counts as one contact, where the reference keeps the rods
 * of its electrode cell layers.  keep[i] = 1 for the rods kept; returns
 * their number. */
uint32_t orc_reduce(const nf_params* p, const float* rods, uint32_t n, const uint32_t* I,
                    const uint32_t* J, const int32_t* image, int64_t nc, double gcut,
                    int32_t min_cont, uint8_t* keep, uint32_t* rounds) {
  const int ax = p->elec_axis;
  const double L = p->box[ax];
  uint32_t* coup = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
  uint32_t* cnt = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = &rods[8 * (size_t)i];
    const double ext = fabs(((double)r[6] * 0.5) * (double)r[3 + ax]);
    const double pos = (double)r[ax];
    coup[i] = (((pos - ext) - (double)r[7]) <= gcut ? 1u : 0u) +
              (((L - (pos + ext)) - (double)r[7]) <= gcut ? 1u : 0u);
    keep[i] = 1;
  }
  uint32_t it = 0, present = n;
  for (;;) {
    ++it;
    for (uint32_t i = 0; i < n; ++i) cnt[i] = coup[i];
    for (int64_t k = 0; k < nc; ++k) {
      int img[3];
      unpack_img(image[k], img);
      if (img[ax] != 0 || !keep[I[k]] || !keep[J[k]]) continue;
      cnt[I[k]]++;
      cnt[J[k]]++;
    }
    uint32_t removed = 0;
    for (uint32_t i = 0; i < n; ++i)
      if (keep[i] && cnt[i] < (uint32_t)min_cont) {
        keep[i] = 0;
        ++removed;
      }
    present -= removed;
    if (removed == 0) break;
  }
  free(coup);
  free(cnt);
  *rounds = it;
  return present;
}

/* Network extraction.  The reference's collect_cunduct_data
 * (src/conductivity.cu:140-423) treats the electrode axis as non-periodic
 * (peri_dir = 7-dir, conductivity.cu:52-77), marks electrode contacts by the
 * buffer layers (inc/conductKernels.h:17-31) and writes COO files for an
 * external Kirchhoff solve (tools/kirchhoff/kirchhoff.py:211-242).  Here the
 * rod graph (contacts whose image does not cross the electrode axis) gets
 * union-find components and an FP64 Laplacian with g = g0 exp(-gap/lambda). */
int64_t orc_network(const nf_params* p, const float* rods, uint32_t n, const uint32_t* I,
                    const uint32_t* J, const double* gap64, const int32_t* image, int64_t nc,
                    uint32_t* labels, uint32_t* row_ptr, uint32_t* col, double* val, int64_t cap,
                    nf_network_info* info) {
  int ax = p->elec_axis;
  if (ax < 0 || ax > 2) return NF_E_INVALID;
  double L = p->box[ax], lam = p->tunnel_lambda, g0 = p->tunnel_g0, gcut = p->contact_gap;
  memset(info, 0, sizeof(*info));
  uint32_t* par = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) par[i] = i;
  uint32_t* deg = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
  uint64_t nedges = 0;
  for (int64_t k = 0; k < nc; ++k) {
    int img[3];
    unpack_img(image[k], img);
    if (img[ax] != 0) continue;
    ++nedges;
    deg[I[k]]++;
    deg[J[k]]++;
    uint32_t a = uf_find(par, I[k]), b = uf_find(par, J[k]);
    if (a != b) {
      if (a < b)
        par[b] = a;
      else
        par[a] = b;
    }
  }
  /* electrode couplings */
  double* gs = (double*)malloc(sizeof(double) * (n ? n : 1));
  double* gk = (double*)malloc(sizeof(double) * (n ? n : 1));
  uint8_t* cs = (uint8_t*)calloc(n ? n : 1, 1);
  uint8_t* ck = (uint8_t*)calloc(n ? n : 1, 1);
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = &rods[8 * (size_t)i];
    double ext = fabs(((double)r[6] * 0.5) * (double)r[3 + ax]);
    double pos = (double)r[ax];
    double gsrc = (pos - ext) - (double)r[7];
    double gsnk = (L - (pos + ext)) - (double)r[7];
    if (gsrc <= gcut) {
      cs[i] = 1;
      gs[i] = g0 * orc_exp(-((gsrc > 0.0 ? gsrc : 0.0) / lam));
      info->n_source++;
    }
    if (gsnk <= gcut) {
      ck[i] = 1;
      gk[i] = g0 * orc_exp(-((gsnk > 0.0 ? gsnk : 0.0) / lam));
      info->n_sink++;
    }
  }
  /* components */
  uint32_t* csize = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
  uint8_t* csrc = (uint8_t*)calloc(n ? n : 1, 1);
  uint8_t* csnk = (uint8_t*)calloc(n ? n : 1, 1);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t r = uf_find(par, i);
    if (labels) labels[i] = r;
    csize[r]++;
    if (cs[i]) csrc[r] = 1;
    if (ck[i]) csnk[r] = 1;
  }
  for (uint32_t i = 0; i < n; ++i)
    if (par[i] == i) {
      info->n_components++;
      if (csize[i] > info->largest) info->largest = csize[i];
      if (csrc[i] && csnk[i]) info->percolating++;
    }
  info->n_edges = nedges;
  info->n_nodes = n + 2;
  /* rows: gather incident entries */
  uint64_t* rs = (uint64_t*)calloc((size_t)n + 1, sizeof(uint64_t));
  for (uint32_t i = 0; i < n; ++i) rs[i + 1] = rs[i] + deg[i];
  ent_t* ents = (ent_t*)malloc(sizeof(ent_t) * (rs[n] ? rs[n] : 1));
  uint64_t* fill = (uint64_t*)calloc((size_t)n + 1, sizeof(uint64_t));
  for (int64_t k = 0; k < nc; ++k) {
    int img[3];
    unpack_img(image[k], img);
    if (img[ax] != 0) continue;
    double gg = gap64[k] > 0.0 ? gap64[k] : 0.0;
    double g = g0 * orc_exp(-(gg / lam));
    uint32_t a = I[k], b = J[k];
    ent_t* e1 = &ents[rs[a] + fill[a]++];
    e1->col = b;
    e1->img = image[k];
    e1->g = g;
    int nimg[3] = {-img[0], -img[1], -img[2]};
    ent_t* e2 = &ents[rs[b] + fill[b]++];
    e2->col = a;
    e2->img = pack_img(nimg);
    e2->g = g;
  }
  int64_t nnz = 0;
  double* tmp = (double*)malloc(sizeof(double) * (n + 2 + 16));
  size_t tmp_cap = n + 2 + 16;
  uint32_t nodes = n + 2;
  int64_t pos_out = 0;
  for (uint32_t r = 0; r < nodes; ++r) {
    if (row_ptr) row_ptr[r] = (uint32_t)pos_out;
    if (r < n) {
      ent_t* e = &ents[rs[r]];
      size_t m = deg[r];
      qsort(e, m, sizeof(ent_t), ent_cmp);
      /* merge duplicate columns, summing in sorted order */
      size_t u = 0;
      uint32_t cols[64];
      (void)cols;
      if (tmp_cap < m + 3) {
        tmp_cap = m + 3;
        tmp = (double*)realloc(tmp, sizeof(double) * tmp_cap);
      }
      uint32_t* ucol = (uint32_t*)malloc(sizeof(uint32_t) * (m + 3));
      for (size_t k = 0; k < m; ++k) {
        if (u > 0 && ucol[u - 1] == e[k].col) {
          tmp[u - 1] = tmp[u - 1] + e[k].g;
        } else {
          ucol[u] = e[k].col;
          tmp[u] = e[k].g;
          ++u;
        }
      }
      if (cs[r]) {
        ucol[u] = n;
        tmp[u] = gs[r];
        ++u;
      }
      if (ck[r]) {
        ucol[u] = n + 1;
        tmp[u] = gk[r];
        ++u;
      }
      double diag = blocked_sum(tmp, u);
      int placed = 0;
      for (size_t k = 0; k <= u; ++k) {
        if (!placed && (k == u || ucol[k] > r)) {
          if (pos_out < cap) {
            if (col) col[pos_out] = r;
            if (val) val[pos_out] = diag;
          }
          ++pos_out;
          placed = 1;
        }
        if (k < u) {
          if (pos_out < cap) {
            if (col) col[pos_out] = ucol[k];
            if (val) val[pos_out] = -tmp[k];
          }
          ++pos_out;
        }
      }
      free(ucol);
    } else {
      const uint8_t* cf = (r == n) ? cs : ck;
      const double* gv = (r == n) ? gs : gk;
      size_t u = 0;
      for (uint32_t i = 0; i < n; ++i)
        if (cf[i]) {
          tmp[u++] = gv[i];
          if (pos_out < cap) {
            if (col) col[pos_out] = i;
            if (val) val[pos_out] = -gv[i];
          }
          ++pos_out;
        }
      double diag = blocked_sum(tmp, u);
      if (pos_out < cap) {
        if (col) col[pos_out] = r;
        if (val) val[pos_out] = diag;
      }
      ++pos_out;
    }
  }
  if (row_ptr) row_ptr[nodes] = (uint32_t)pos_out;
  nnz = pos_out;
  info->nnz = (uint64_t)nnz;
  free(par);
  free(deg);
  free(gs);
  free(gk);
  free(cs);
  free(ck);
  free(csize);
  free(csrc);
  free(csnk);
  free(rs);
  free(ents);
  free(fill);
  free(tmp);
  return nnz;
}

/* ===================================================================== */
/* atan, van der Waals energy, centre-centre distance statistics          */
/* ===================================================================== */
/* atan from +, -, *, / and the correctly rounded C sqrt: two half-angle
 * reductions to |a| <= tan(pi/16), the alternating series to a^35, and
 * pi/2 - atan(1/x) above 1 (the library evaluates the same operations). */
double orc_atan(double x) {
  if (x != x) return x;
  int neg = x < 0.0;
  double a = neg ? -x : x;
  int inv = a > 1.0;
  if (inv) a = 1.0 / a;
  a = a / (1.0 + sqrt(1.0 + a * a));
  a = a / (1.0 + sqrt(1.0 + a * a));
  double t = a * a;
  double s = 1.0 / 35.0;
  static const double den[] = {33, 31, 29, 27, 25, 23, 21, 19, 17, 15, 13, 11, 9, 7, 5, 3, 1};
  for (int k = 0; k < 17; ++k) s = 1.0 / den[k] - t * s;
  double r = 4.0 * (a * s);
  if (inv) r = 1.57079632679489661923 - r;
  return neg ? -r : r;
}

/* One pair of the van der Waals model of statKernel_VdW_naive
 * (inc/statKernel_MUT.h:531-628; exclArea_naive :800-842 with the second
 * radius in the skew branch as exclArea_skew :869-871; curve2surface_limit
 * :891-894), rod 2 shifted by sh.  Returns 1 and the energy when the pair's
 * closest points are interior to both rods and its gap is below dlim. */
int orc_vdw_pair(const float* r1, const float* r2, const double sh[3], double dlim, double A,
                 double B, double* energy) {
  const double PI = 3.14159265358979323846;
  double c1[3] = {r1[0], r1[1], r1[2]}, c2[3], u1[3] = {r1[3], r1[4], r1[5]},
         u2[3] = {r2[3], r2[4], r2[5]};
  for (int d = 0; d < 3; ++d) c2[d] = (double)r2[d] + sh[d];
  double h1 = (double)(0.5f * r1[6]), h2 = (double)(0.5f * r2[6]), a1 = r1[7], a2 = r2[7];
  double w[3] = {c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]};
  double aa = u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2];
  double ee = u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2];
  double bb = u1[0] * u2[0] + u1[1] * u2[1] + u1[2] * u2[2];
  double cg = bb / sqrt(aa * ee);
  double sg2 = 1.0 - cg * cg;
  double sg = sg2 > 0.0 ? sqrt(sg2) : 0.0;
  const double eps = 1e-6;
  double excA = 0.0, dline;
  if (sg > eps) {
    double d1 = u1[0] * w[0] + u1[1] * w[1] + u1[2] * w[2];
    double d2 = u2[0] * w[0] + u2[1] * w[1] + u2[2] * w[2];
    double den = aa * ee - bb * bb;
    double t1 = (bb * d2 - ee * d1) / den;
    double t2 = (aa * d2 - bb * d1) / den;
    if (!(fabs(t1) <= h1 && fabs(t2) <= h2)) return 0;
    double p[3];
    for (int d = 0; d < 3; ++d) p[d] = (w[d] + t1 * u1[d]) - t2 * u2[d];
    dline = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    double ea = 2.0 * a1 / sg, eb = 2.0 * a2 / sg;
    excA = ea * eb * sg;
    double cx = h1 - fabs(t1), cy = h2 - fabs(t2);
    if (fabs(cx) < ea || fabs(cy) < eb) excA = excA / 2.0;
    if (cx <= -ea || cy <= -eb) excA = 0.0;
  } else {
    double n1 = sqrt(aa);
    double pr = (-(u1[0] * w[0] + u1[1] * w[1] + u1[2] * w[2])) / n1;
    double q[3];
    for (int d = 0; d < 3; ++d) q[d] = -w[d] - pr * u1[d] / n1;
    dline = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    double tip0 = fabs(pr) - h2, tip1 = tip0 + 2.0 * h2;
    if (tip1 < -h1 || tip0 > h1) return 0;
    if (tip0 <= -h1 + eps && tip1 >= h1 - eps) excA = 2.0 * a1 * (2.0 * h1);
    if (tip0 > h1 - eps) excA = 0.0;
    if (tip0 > -h1 + eps && tip1 < h1 - eps) excA = 2.0 * 2.0 * a1 * ((tip1 < h1 ? tip1 : h1) - tip0);
  }
  double gap = (dline - a1) - a2;
  if (!(gap < dlim)) return 0;
  double theta = orc_atan(a1 / ((dline - a2) - dlim));
  double d = gap * 1e4, dl = dlim * 1e4, R1 = a1 * 1e4, R2 = a2 * 1e4;
  double qq = 1.0 - theta * d / dl / dl / (R1 / R2 + 1.0);
  double d4 = (d * d) * (d * d), dl4 = (dl * dl) * (dl * dl);
  double d10 = d4 * d4 * (d * d), dl10 = dl4 * dl4 * (dl * dl);
  double q4 = (qq * qq) * (qq * qq);
  double c2s = -A * PI * R1 / 2.0 * theta * (1.0 / d4 - 1.0 / dl4 * (1.0 / qq)) +
               B * 2.0 * PI * R1 / 5.0 * theta * (1.0 / d10 - 1.0 / dl10 * (1.0 / q4));
  *energy = excA / theta / a1 * c2s;
  return 1;
}

/* Van der Waals energy over a contact list sorted by (I, J, image): per-row
 * sums in row order, rows summed in blocks of 1024 in order, block sums in
 * order.  Returns the contributing pair count. */
uint64_t orc_vdw(const nf_params* p, const float* rods, uint32_t n, const uint32_t* I,
                 const uint32_t* J, const int32_t* image, int64_t nc, double dlim, double A,
                 double B, double* energy) {
  double* rs = (double*)calloc(n ? n : 1, sizeof(double));
  uint64_t cnt = 0;
  for (int64_t k = 0; k < nc; ++k) {
    int img[3];
    unpack_img(image[k], img);
    double sh[3] = {img[0] * p->box[0], img[1] * p->box[1], img[2] * p->box[2]};
    double e;
    if (orc_vdw_pair(&rods[8 * (size_t)I[k]], &rods[8 * (size_t)J[k]], sh, dlim, A, B, &e)) {
      rs[I[k]] = rs[I[k]] + e;
      ++cnt;
    }
  }
  double tot = 0.0;
  for (uint32_t b = 0; b < n; b += 1024) {
    double s = 0.0;
    uint32_t e = b + 1024 < n ? b + 1024 : n;
    for (uint32_t r = b; r < e; ++r) s = s + rs[r];
    tot = tot + s;
  }
  free(rs);
  *energy = tot;
  return cnt;
}

/* Centre-centre distance statistics in the manner of statKernel_CENT
 * (inc/statKernel_CENT.h:8-122): every unordered pair of distinct rods, over
 * all periodic images, whose centre distance d is below bins * width;
 * counts[k] for d in [k width, (k+1) width) (decided by comparing d^2 with
 * the squared edges), and exact fixed-point sums of round(d * 2^24) and
 * round(d^2 * 2^16).  d is evaluated with the lower id first and the image
 * applied to the higher id.  Brute force over all pairs. */
uint64_t orc_centre_distance(const nf_params* p, const float* rods, uint32_t n, int32_t bins,
                             double width, int64_t* counts, uint64_t* sum_q, uint64_t* sum2_q) {
  double rmax = bins * width;
  double* e2 = (double*)malloc(sizeof(double) * (size_t)(bins + 1));
  for (int k = 0; k <= bins; ++k) e2[k] = ((double)k * width) * ((double)k * width);
  int nim[3];
  for (int d = 0; d < 3; ++d) nim[d] = (int)ceil(rmax / p->box[d]) + 1;
  memset(counts, 0, sizeof(int64_t) * (size_t)bins);
  uint64_t cnt = 0, sq = 0, s2q = 0;
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = i + 1; j < n; ++j)
      for (int iz = -nim[2]; iz <= nim[2]; ++iz)
        for (int iy = -nim[1]; iy <= nim[1]; ++iy)
          for (int ix = -nim[0]; ix <= nim[0]; ++ix) {
            const float* a = &rods[8 * (size_t)i];
            const float* b = &rods[8 * (size_t)j];
            double dx = ((double)a[0] - (double)b[0]) - ix * p->box[0];
            double dy = ((double)a[1] - (double)b[1]) - iy * p->box[1];
            double dz = ((double)a[2] - (double)b[2]) - iz * p->box[2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (!(d2 < e2[bins])) continue;
            int lo = 0, hi = bins; /* largest k with e2[k] <= d2 */
            while (hi - lo > 1) {
              int mid = (lo + hi) / 2;
              if (e2[mid] <= d2) lo = mid; else hi = mid;
            }
            counts[lo]++;
            ++cnt;
            double d = sqrt(d2);
            sq += (uint64_t)floor(d * 16777216.0 + 0.5);
            s2q += (uint64_t)floor(d2 * 65536.0 + 0.5);
          }
  free(e2);
  *sum_q = sq;
  *sum2_q = s2q;
  return cnt;
}

/* ===================================================================== */
/* Node-level segment network: contact points and screening              */
/* ===================================================================== */
/* FP64 squared distance between two segments (centre, half vector), the
 * clamped closest-point solve of orc_pair_d2; a zero-length first segment
 * is a point. */
static double orc_seg_seg_d2(const double c1[3], const double h1[3], const double c2[3],
                             const double h2[3]) {
  double wx = c1[0] - c2[0], wy = c1[1] - c2[1], wz = c1[2] - c2[2];
  double a = (h1[0] * h1[0] + h1[1] * h1[1]) + h1[2] * h1[2];
  double e = (h2[0] * h2[0] + h2[1] * h2[1]) + h2[2] * h2[2];
  double b = (h1[0] * h2[0] + h1[1] * h2[1]) + h1[2] * h2[2];
  double c = (h1[0] * wx + h1[1] * wy) + h1[2] * wz;
  double f = (h2[0] * wx + h2[1] * wy) + h2[2] * wz;
  double denom = a * e - b * b;
  double s = 0.0;
  if (denom > 0.0) {
    s = (b * f - c * e) / denom;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
  }
  double t = (b * s + f) / e;
  if (t < -1.0) {
    t = -1.0;
    s = a > 0.0 ? (-b - c) / a : 0.0;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
  } else if (t > 1.0) {
    t = 1.0;
    s = a > 0.0 ? (b - c) / a : 0.0;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
  }
  double vx = (wx + s * h1[0]) - t * h2[0];
  double vy = (wy + s * h1[1]) - t * h2[1];
  double vz = (wz + s * h1[2]) - t * h2[2];
  return (vx * vx + vy * vy) + vz * vz;
}

/* Closest-point parameters of one contact (I, J, image on J) in length
 * units, as the closest points of get_contact_arrays' (I, J, tI, tJ)
 * records (src/get_contact_arrays.cu:71-140). */
static void orc_contact_params(const nf_params* p, const float* ri, const float* rj, int32_t code,
                               double* tI, double* tJ) {
  double sh[3];
  for (int d = 0; d < 3; ++d) sh[d] = (double)(int8_t)((code >> (8 * d)) & 0xFF) * p->box[d];
  double wx = ((double)ri[0] - (double)rj[0]) - sh[0];
  double wy = ((double)ri[1] - (double)rj[1]) - sh[1];
  double wz = ((double)ri[2] - (double)rj[2]) - sh[2];
  double hli = (double)ri[6] * 0.5, hlj = (double)rj[6] * 0.5;
  double h1x = hli * (double)ri[3], h1y = hli * (double)ri[4], h1z = hli * (double)ri[5];
  double h2x = hlj * (double)rj[3], h2y = hlj * (double)rj[4], h2z = hlj * (double)rj[5];
  double a = (h1x * h1x + h1y * h1y) + h1z * h1z;
  double e = (h2x * h2x + h2y * h2y) + h2z * h2z;
  double b = (h1x * h2x + h1y * h2y) + h1z * h2z;
  double c = (h1x * wx + h1y * wy) + h1z * wz;
  double f = (h2x * wx + h2y * wy) + h2z * wz;
  double denom = a * e - b * b;
  double s = 0.0;
  if (denom > 0.0) {
    s = (b * f - c * e) / denom;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
  }
  double t = (b * s + f) / e;
  if (t < -1.0) {
    t = -1.0;
    s = (-b - c) / a;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
  } else if (t > 1.0) {
    t = 1.0;
    s = (b - c) / a;
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
  }
  *tI = s * hli;
  *tJ = t * hlj;
}

/* Contact points and screening of every contact (flag 0 kept, 1 screened,
 * 2 crossing the electrode axis): the segment between the contact points,
 * with the radius min(a_I, a_J) of generate_virt_inclusions
 * (inc/conductKernels.h:289-361), is screened by any periodic image of any
 * third rod K whose axis comes closer than that radius + a_K
 * (check_screened_contacts, src/conductivity.cu:430-528).  Brute force over
 * all rods and the images within reach. */
void orc_node_points(const nf_params* p, const float* rods, uint32_t n, const uint32_t* I,
                     const uint32_t* J, const int32_t* img, uint64_t nc, double* tI, double* tJ,
                     uint32_t* flag, int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  const int ax = p->elec_axis;
  double hlmax = 0.0, amax = 0.0;
  for (uint32_t k = 0; k < n; ++k) {
    if ((double)rods[8 * (size_t)k + 6] * 0.5 > hlmax) hlmax = (double)rods[8 * (size_t)k + 6] * 0.5;
    if ((double)rods[8 * (size_t)k + 7] > amax) amax = (double)rods[8 * (size_t)k + 7];
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < (int64_t)nc; ++q) {
    const float* ri = rods + 8 * (size_t)I[q];
    const float* rj = rods + 8 * (size_t)J[q];
    const int32_t code = img[q];
    if ((int)(int8_t)((code >> (8 * ax)) & 0xFF) != 0) {
      flag[q] = 2;
      tI[q] = tJ[q] = 0.0;
      continue;
    }
    double ti, tj;
    orc_contact_params(p, ri, rj, code, &ti, &tj);
    tI[q] = ti;
    tJ[q] = tj;
    double u[3];
    for (int d = 0; d < 3; ++d) u[d] = (double)(int8_t)((code >> (8 * d)) & 0xFF) * p->box[d];
    double A[3], B[3];
    for (int d = 0; d < 3; ++d) {
      A[d] = (double)ri[d] + ti * (double)ri[3 + d];
      B[d] = ((double)rj[d] + u[d]) + tj * (double)rj[3 + d];
    }
    double av = (double)ri[7] < (double)rj[7] ? (double)ri[7] : (double)rj[7];
    double m[3] = {0.5 * (A[0] + B[0]), 0.5 * (A[1] + B[1]), 0.5 * (A[2] + B[2])};
    double hv[3] = {0.5 * (B[0] - A[0]), 0.5 * (B[1] - A[1]), 0.5 * (B[2] - A[2])};
    double hs = sqrt((hv[0] * hv[0] + hv[1] * hv[1]) + hv[2] * hv[2]);
    double reach = hlmax + hs + av + amax;
    int K[3];
    for (int d = 0; d < 3; ++d) K[d] = (int)floor(reach / p->box[d]) + 2;
    int scr = 0;
    for (uint32_t k = 0; k < n && !scr; ++k) {
      if (k == I[q] || k == J[q]) continue;
      const float* rk = rods + 8 * (size_t)k;
      double hl = (double)rk[6] * 0.5;
      double h2[3] = {hl * (double)rk[3], hl * (double)rk[4], hl * (double)rk[5]};
      double R = av + (double)rk[7];
      /* images of K around the midpoint: start from the nearest one */
      int n0[3];
      for (int d = 0; d < 3; ++d) n0[d] = (int)floor((m[d] - (double)rk[d]) / p->box[d] + 0.5);
      for (int iz = n0[2] - K[2]; iz <= n0[2] + K[2] && !scr; ++iz)
        for (int iy = n0[1] - K[1]; iy <= n0[1] + K[1] && !scr; ++iy)
          for (int ix = n0[0] - K[0]; ix <= n0[0] + K[0] && !scr; ++ix) {
            double c2[3] = {(double)rk[0] + (double)ix * p->box[0], (double)rk[1] + (double)iy * p->box[1],
                            (double)rk[2] + (double)iz * p->box[2]};
            double d2 = orc_seg_seg_d2(m, hv, c2, h2);
            if (d2 < R * R) scr = 1;
          }
    }
    flag[q] = scr ? 1u : 0u;
  }
}

/* ===================================================================== */
/* expand                                                                  */
/* ===================================================================== */
int64_t orc_expand(const float* rods, uint32_t n, int k, double L, double B, float* out, int64_t cap) {
  int64_t nb = 0, na = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const double x = (double)rods[8 * (size_t)i + k];
    if (x >= L - B) ++nb;
    if (x < B) ++na;
  }
  const int64_t total = (int64_t)n + nb + na;
  if (!out || cap < total) return total;
  int64_t ib = n, ia = n + nb;
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = rods + 8 * (size_t)i;
    const double x = (double)r[k];
    float* o = out + 8 * (size_t)i;
    memcpy(o, r, 8 * sizeof(float));
    o[k] = (float)(x + B);
    if (x >= L - B) {
      o = out + 8 * (size_t)ib++;
      memcpy(o, r, 8 * sizeof(float));
      o[k] = (float)(x + (B - L));
    }
    if (x < B) {
      o = out + 8 * (size_t)ia++;
      memcpy(o, r, 8 * sizeof(float));
      o[k] = (float)(x + (B + L));
    }
  }
  return total;
}
