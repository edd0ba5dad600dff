// nf_common.h - arithmetic shared by the HIP kernels and the host side of
// the library: Philox4x32-10, uniform conversion, exp/log from basic IEEE
// operations, candidate sampling, and the capsule (segment-segment) distance
// predicates in FP64 and in FP32 with a rigorous error band.
//
// Every double expression is written in the exact operation order the
// candidate stream and the FP64 decisions are defined by; the library is
// compiled with -ffp-contract=off so no operation is fused or reordered.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nanofiller.h"

#define NF_HD __host__ __device__ __forceinline__

namespace nf {

// --------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter-based, so a
// candidate's random numbers depend only on (slot, round, draw, tag) and the
// key (seed, stream); replaces curand XORWOW seeded from clock()
// (src/simPrms.cpp:415, inc/cuda_CNTgenerate.h:200).
// --------------------------------------------------------------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

NF_HD u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return u32x4{c0, c1, c2, c3};
}

// 52 random bits -> double in (0, 1); exact (m + 0.5 < 2^52 fits 53 bits).
NF_HD double u01(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)(a >> 12) << 32) | (uint64_t)b;
  return ((double)m + 0.5) * 0x1p-52;
}

// --------------------------------------------------------------------------
// exp / log from +, -, *, / only (Cody-Waite reduction; Taylor series for
// e^r, |r| <= ln2/2; atanh series for log of m in [sqrt(1/2), sqrt(2))).
// Bit-identical on the CPU and the GPU because each step is one correctly
// rounded IEEE operation.
// --------------------------------------------------------------------------
NF_HD double nf_bits_to_double(uint64_t u) { return __builtin_bit_cast(double, u); }

// Correctly rounded square root.  The hardware / library sqrt of the GPU is
// not guaranteed to round correctly in FP64, so: estimate, one Newton step
// with fused residual, then an exact residual test picks between the
// neighbouring doubles (for y = k*ulp the residual x - y^2 is an integer
// multiple of ulp^2, so "x above the upper midpoint" is exactly
// r > y*ulp and "x below the lower midpoint" is r <= -y*ulp_below).
NF_HD double sqrt_cr(double x) {
  if (!(x > 0.0) || !(x < 0x1p1023)) return sqrt(x);
  double scale = 1.0;
  if (x < 0x1p-900) {
    x = x * 0x1p1000;
    scale = 0x1p-500;
  } else if (x > 0x1p900) {
    x = x * 0x1p-1000;
    scale = 0x1p500;
  }
  double y = sqrt(x);
  double r = fma(-y, y, x);
  y = fma(r, 0.5 / y, y);
  r = fma(-y, y, x);
  const uint64_t b = __builtin_bit_cast(uint64_t, y);
  const double u_up = __builtin_bit_cast(double, ((b >> 52) - 52ull) << 52);
  const double u_dn = (b & 0x000FFFFFFFFFFFFFull) == 0 ? u_up * 0.5 : u_up;
  if (r > y * u_up)
    y = y + u_up;
  else if (r <= -(y * u_dn))
    y = y - u_dn;
  return y * scale;
}

NF_HD double pow2i(int k) {
  return nf_bits_to_double((uint64_t)(k + 1023) << 52);
}

NF_HD double dexp(double x) {
  if (x != x) return x;
  if (x > 709.0) return nf_bits_to_double(0x7FF0000000000000ull);
  if (x < -745.0) return 0.0;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  const double INV_LN2 = 1.44269504088896338700e+00;
  const double kd = floor(x * INV_LN2 + 0.5);
  const int k = (int)kd;
  const double r = (x - kd * LN2_HI) - kd * LN2_LO;
  double p = 1.0 / 6227020800.0;
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

NF_HD double dlog(double x) {
  if (x != x || x < 0.0) return nf_bits_to_double(0x7FF8000000000000ull);
  if (x == 0.0) return nf_bits_to_double(0xFFF0000000000000ull);
  if (x == nf_bits_to_double(0x7FF0000000000000ull)) return x;
  int e = 0;
  if (x < 0x1p-1022) {
    x = x * 0x1p54;
    e = -54;
  }
  uint64_t u = __builtin_bit_cast(uint64_t, x);
  e += (int)((u >> 52) & 0x7FF) - 1023;
  u = (u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
  double m = nf_bits_to_double(u);
  if (m > 1.41421356237309504880) {
    m = m * 0.5;
    e += 1;
  }
  const double f = (m - 1.0) / (m + 1.0);
  const double f2 = f * f;
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
  const double lm = (2.0 * f) * s;
  const double ed = (double)e;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  return ed * LN2_HI + (lm + ed * LN2_LO);
}

// --------------------------------------------------------------------------
// Per-instance sampling description (device-resident, one per packing).
// --------------------------------------------------------------------------
struct SampleDesc {
  double box[3];
  double len, len_sigma, len_min, len_max;
  double radius, sep, soft_frac;
  double kappa;
  double axis[3], e1[3], e2[3];
  double agg_sigma, agg_peak;
  uint32_t seed, stream;
  int32_t len_law, ort_law, den_law, n_agg;
  const double* agg;  // 3 doubles per agglomerate centre
  const float* mesh;  // density map, C order (x slowest), NF_DEN_MESH
  int32_t mext[3];
};

// Density map value at (x, y, z): trilinear interpolation with periodic
// wrap, mesh points at the centres of the map's cells - the reference's
// normalised, linearly filtered, wrapping 3-D texture (src/den2tex.h:5-53,
// read at inc/cuda_CNTgenerate.h:228), in FP64 with a fixed evaluation
// order (z, then y, then x).
NF_HD double mesh_density(const float* v, const int32_t e[3], const double box[3], double x, double y,
                          double z) {
  const double pos[3] = {x, y, z};
  int i0[3], i1[3];
  double f[3];
  for (int a = 0; a < 3; ++a) {
    const double t = (pos[a] / box[a]) * (double)e[a] - 0.5;
    const double fl = floor(t);
    f[a] = t - fl;
    int k = (int)fl;  // -1 .. e - 1 for positions in [0, box)
    if (k < 0) k += e[a];
    if (k >= e[a]) k -= e[a];
    i0[a] = k;
    i1[a] = k + 1 == e[a] ? 0 : k + 1;
  }
  const uint64_t sy = (uint64_t)e[2], sx = (uint64_t)e[1] * (uint64_t)e[2];
  const uint64_t x0 = (uint64_t)i0[0] * sx, x1 = (uint64_t)i1[0] * sx;
  const uint64_t y0 = (uint64_t)i0[1] * sy, y1 = (uint64_t)i1[1] * sy;
  const uint64_t z0 = (uint64_t)i0[2], z1 = (uint64_t)i1[2];
  const double gz = 1.0 - f[2], gy = 1.0 - f[1], gx = 1.0 - f[0];
  const double c00 = (double)v[x0 + y0 + z0] * gz + (double)v[x0 + y0 + z1] * f[2];
  const double c01 = (double)v[x0 + y1 + z0] * gz + (double)v[x0 + y1 + z1] * f[2];
  const double c10 = (double)v[x1 + y0 + z0] * gz + (double)v[x1 + y0 + z1] * f[2];
  const double c11 = (double)v[x1 + y1 + z0] * gz + (double)v[x1 + y1 + z1] * f[2];
  const double c0 = c00 * gy + c01 * f[1];
  const double c1 = c10 * gy + c11 * f[1];
  return c0 * gx + c1 * f[0];
}

// Basis perpendicular to the vMF axis mu: the least-aligned coordinate axis
// projected out of mu, normalised, and mu x e1.
NF_HD void vmf_basis(const double mu[3], double e1[3], double e2[3]) {
  double t[3] = {0.0, 0.0, 0.0};
  const double ax = fabs(mu[0]), ay = fabs(mu[1]), az = fabs(mu[2]);
  if (ax <= ay && ax <= az)
    t[0] = 1.0;
  else if (ay <= az)
    t[1] = 1.0;
  else
    t[2] = 1.0;
  const double d = (t[0] * mu[0] + t[1] * mu[1]) + t[2] * mu[2];
  const double u0 = t[0] - d * mu[0], u1 = t[1] - d * mu[1], u2 = t[2] - d * mu[2];
  const double nrm = sqrt_cr((u0 * u0 + u1 * u1) + u2 * u2);
  e1[0] = u0 / nrm;
  e1[1] = u1 / nrm;
  e1[2] = u2 / nrm;
  e2[0] = mu[1] * e1[2] - mu[2] * e1[1];
  e2[1] = mu[2] * e1[0] - mu[0] * e1[2];
  e2[2] = mu[0] * e1[1] - mu[1] * e1[0];
}

// Agglomerate centre k: Philox counter (k, 0, d, 1).
NF_HD void agg_centre(uint32_t seed, uint32_t stream, const double box[3], uint32_t k,
                      double out[3]) {
  const u32x4 o = philox(k, 0u, 0u, 1u, seed, stream);
  const u32x4 q = philox(k, 0u, 1u, 1u, seed, stream);
  out[0] = u01(o.x, o.y) * box[0];
  out[1] = u01(o.z, o.w) * box[1];
  out[2] = u01(q.x, q.y) * box[2];
}

// Rod record as the kernels store it: P = (x, y, z, l/2), D = (cx, cy, cz, a).
struct Rod {
  float4 P, D;
};

// --------------------------------------------------------------------------
// Capsule distance, FP64.  Axis segments P = r1 + s h1, Q = r2 + shift +
// t h2 with s, t in [-1, 1] and h = (l/2) c; the exact minimiser of |P-Q|^2
// (clamp s, solve t, clamp t and re-solve s).  Plays the role of
// distance_cylinder_capped (inc/intersect/intersect_cylinder_capped.h:69-147)
// and cnt_intersec_shr (inc/intersection_dev.h:10-80), without their l/2+a
// clamp slack and single-pass clamp.  *within: centre distance within
// (l1/2 + l2/2) * REACH_PAD + R.
// --------------------------------------------------------------------------
constexpr double REACH_PAD = 1.000001;

// As pair_d2_f64, also returning the closest points' axial parameters in
// units of the half lengths (s on rod 1, t on rod 2, both in [-1, 1]).
NF_HD double pair_closest_f64(const float4 P1, const float4 D1, const float4 P2, const float4 D2,
                              const double sx, const double sy, const double sz, const double R,
                              bool* within, double* s_out, double* t_out) {
  const double wx = ((double)P1.x - (double)P2.x) - sx;
  const double wy = ((double)P1.y - (double)P2.y) - sy;
  const double wz = ((double)P1.z - (double)P2.z) - sz;
  const double hli = (double)P1.w, hlj = (double)P2.w;
  const double rr = (hli + hlj) * REACH_PAD + R;
  const double w2 = (wx * wx + wy * wy) + wz * wz;
  *within = (w2 <= rr * rr);
  const double h1x = hli * (double)D1.x, h1y = hli * (double)D1.y, h1z = hli * (double)D1.z;
  const double h2x = hlj * (double)D2.x, h2y = hlj * (double)D2.y, h2z = hlj * (double)D2.z;
  const double a = (h1x * h1x + h1y * h1y) + h1z * h1z;
  const double e = (h2x * h2x + h2y * h2y) + h2z * h2z;
  const double b = (h1x * h2x + h1y * h2y) + h1z * h2z;
  const double c = (h1x * wx + h1y * wy) + h1z * wz;
  const double f = (h2x * wx + h2y * wy) + h2z * wz;
  const double denom = a * e - b * b;
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
  const double vx = (wx + s * h1x) - t * h2x;
  const double vy = (wy + s * h1y) - t * h2y;
  const double vz = (wz + s * h1z) - t * h2z;
  if (s_out) *s_out = s;
  if (t_out) *t_out = t;
  return (vx * vx + vy * vy) + vz * vz;
}

NF_HD double pair_d2_f64(const float4 P1, const float4 D1, const float4 P2, const float4 D2,
                         const double sx, const double sy, const double sz, const double R,
                         bool* within) {
  return pair_closest_f64(P1, D1, P2, D2, sx, sy, sz, R, within, nullptr, nullptr);
}

// Squared distance between two FP64 segments given by centre and half
// vector, with the clamped closest-point solve of pair_closest_f64; a
// zero-length first segment is a point.
NF_HD double seg_seg_d2(double c1x, double c1y, double c1z, double h1x, double h1y, double h1z,
                        double c2x, double c2y, double c2z, double h2x, double h2y, double h2z) {
  const double wx = c1x - c2x, wy = c1y - c2y, wz = c1z - c2z;
  const double a = (h1x * h1x + h1y * h1y) + h1z * h1z;
  const double e = (h2x * h2x + h2y * h2y) + h2z * h2z;
  const double b = (h1x * h2x + h1y * h2y) + h1z * h2z;
  const double c = (h1x * wx + h1y * wy) + h1z * wz;
  const double f = (h2x * wx + h2y * wy) + h2z * wz;
  const double denom = a * e - b * b;
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
  const double vx = (wx + s * h1x) - t * h2x;
  const double vy = (wy + s * h1y) - t * h2y;
  const double vz = (wz + s * h1z) - t * h2z;
  return (vx * vx + vy * vy) + vz * vz;
}

// --------------------------------------------------------------------------
// Contact (tunnelling) area S of a rod pair: the model of
// surf_to_cylinder_capped (inc/intersect/intersect_cylinder_capped.h:204-293)
// fed with the line geometry of distance_cylinder_capped (:69-147), as
// get_surface_cyl2cyl computes it per contact (src/get_contact_data.cu:
// 130-172), evaluated in FP64.  Rod 2 sits at its periodic image
// (P2 + shift).  Directions are unit vectors, t in length units.
//
//  - Lines: closest points of the two infinite axes (t1, t2) and their
//    distance D; axes within 1e-6 of parallel take t1 = +-l1/2 towards rod 2
//    and t2 its projection (the reference's parallel branch).
//  - Point contact: the disc of radius dr = min(dr1, dr2) with
//    dr_i = a_i sqrt(1 - (h_i / a_i)^2), h1 = max(D - a2 - rlim, 0),
//    h2 = max(D - a1 - rlim, 0) (dr_i = 0 when h_i > a_i); S = pi dr^2.
//  - Crossing (both closest points at least a radius inside their rods): an
//    ellipse of semi-axes dr and (A+ + A-)/2 along the bisector ax1 of c1 and
//    c2' = c2 sign(c1.c2); A+- is the smallest of dr2 / sin(gamma/2) (the
//    strip-overlap extent) and the distances along ax1 from the contact
//    points to the rod ends on that side, at least dr:
//    S = max(S, pi/2 dr (A+ + A-)).
//  - Near-parallel (sin gamma <= max(a) / min(l)): the overlap of the two
//    rods' projections on c1 times the strip width 2 dr, as the reference's
//    end-point projections (a separation rather than an overlap when the
//    projections do not overlap): S = max(S, |p2 - p1| 2 dr).
//
// Two terms differ from the reference on purpose: the rod-end distances are
// measured from the contact point (the reference adds the rod's centre
// position r1 / r2 to them, :257-260, which makes S depend on where in the
// box the pair lies), and rod 2's contact parameter along c2' is t2 sign(c1.c2)
// (the reference's copysign(t2, c1.c2) drops the sign of t2, :253).
// 1 - d^2 is clamped at 0 before its square root.
// --------------------------------------------------------------------------
NF_HD double dmin2(double a, double b) { return b < a ? b : a; }
NF_HD double dmax2(double a, double b) { return b > a ? b : a; }

NF_HD double contact_area_f64(const float4 P1, const float4 D1, const float4 P2, const float4 D2,
                              const double sx, const double sy, const double sz, const double rlim) {
  const double PI = 3.14159265358979323846;
  const double l1 = (double)P1.w * 2.0, l2 = (double)P2.w * 2.0;
  const double a1 = (double)D1.w, a2 = (double)D2.w;
  const double c1x = (double)D1.x, c1y = (double)D1.y, c1z = (double)D1.z;
  const double c2x = (double)D2.x, c2y = (double)D2.y, c2z = (double)D2.z;
  const double bx = ((double)P2.x - (double)P1.x) + sx;
  const double by = ((double)P2.y - (double)P1.y) + sy;
  const double bz = ((double)P2.z - (double)P1.z) + sz;
  const double d = (c1x * c2x + c1y * c2y) + c1z * c2z;
  const double bc1 = (bx * c1x + by * c1y) + bz * c1z;
  const double bc2 = (bx * c2x + by * c2y) + bz * c2z;
  double t1, t2;
  if (1.0 - fabs(d) > 1e-6) {
    const double den = d * d - 1.0;
    t1 = (d * bc2 - bc1) / den;
    t2 = (bc2 - d * bc1) / den;
  } else {
    t1 = bc1 < 0.0 ? -(l1 * 0.5) : l1 * 0.5;
    t2 = d * t1 - bc2;
  }
  const double vx = (t1 * c1x - t2 * c2x) - bx;
  const double vy = (t1 * c1y - t2 * c2y) - by;
  const double vz = (t1 * c1z - t2 * c2z) - bz;
  const double D = sqrt_cr((vx * vx + vy * vy) + vz * vz);
  const double h1 = dmax2((D - a2) - rlim, 0.0), h2 = dmax2((D - a1) - rlim, 0.0);
  double dr1 = 0.0, dr2 = 0.0;
  if (h1 <= a1) {
    const double q = h1 / a1;
    dr1 = a1 * sqrt_cr(dmax2(1.0 - q * q, 0.0));
  }
  if (h2 <= a2) {
    const double q = h2 / a2;
    dr2 = a2 * sqrt_cr(dmax2(1.0 - q * q, 0.0));
  }
  const double dr = dmin2(dr1, dr2);
  double S = (PI * dr) * dr;
  const double sg = d < 0.0 ? -1.0 : 1.0;
  const double dp = d * sg;  // c1 . c2'
  const double sng = sqrt_cr(dmax2(1.0 - dp * dp, 0.0));
  const double at1 = fabs(t1), at2 = fabs(t2);
  if (at1 < l1 * 0.5 - a1 && at2 < l2 * 0.5 - a2) {
    // bisector of c1 and c2'
    double ax = c1x + sg * c2x, ay = c1y + sg * c2y, az = c1z + sg * c2z;
    const double an = sqrt_cr((ax * ax + ay * ay) + az * az);
    ax = ax / an;
    ay = ay / an;
    az = az / an;
    const double csg = sqrt_cr(dmax2(1.0 - sng * sng, 0.0));
    const double sngh = sqrt_cr((1.0 - csg) * 0.5);
    const double t2p = sg * t2;
    const double k1 = (c1x * ax + c1y * ay) + c1z * az;                   // c1 . ax1
    const double k2 = ((sg * c2x) * ax + (sg * c2y) * ay) + (sg * c2z) * az;  // c2' . ax1
    const double dt0 = sngh > 1e-6 ? dr2 / sngh : l1 * 0.5;
    const double dt1p = (l1 * 0.5 - t1) * k1;
    const double dt1m = fabs((-(l1 * 0.5) - t1) * k1);
    const double dt2p = (l2 * 0.5 - t2p) * k2;
    const double dt2m = fabs((-(l2 * 0.5) - t2p) * k2);
    const double Ap = dmax2(dr, dmin2(dt0, dmin2(dt1p, dt2p)));
    const double Am = dmax2(dr, dmin2(dt0, dmin2(dt1m, dt2m)));
    S = dmax2(S, ((0.5 * PI) * dr) * (Ap + Am));
  }
  const double amax = dmax2(a1, a2), lmin = dmin2(l1, l2);
  if (sng <= amax / lmin && (at1 <= l1 * 0.5 + a1 || at2 <= l2 * 0.5 + a2)) {
    // rod ends projected on c1, relative to rod 1's centre
    const double t1a = bc1 < 0.0 ? -(l1 * 0.5) : l1 * 0.5;
    const double t2a = bc2 < 0.0 ? -(l2 * 0.5) : l2 * 0.5;
    const double p1a = t1a, p1b = -t1a;  // (c1 t) . c1 with |c1| = 1
    const double p2a = bc1 + t2a * d, p2b = bc1 - t2a * d;
    const double p1 = bc1 > 0.0 ? dmin2(p1a, p2a) : dmax2(p1a, p2a);
    const double p2 = bc1 > 0.0 ? dmax2(p1b, p2b) : dmin2(p1b, p2b);
    S = dmax2(S, (fabs(p2 - p1) * 2.0) * dr);
  }
  return S;
}

// Deterministic atan from +, -, *, / and the correctly rounded sqrt: two
// half-angle reductions atan(a) = 2 atan(a / (1 + sqrt(1 + a^2))) bring
// |a| <= tan(pi/16), then the alternating series to a^35; |x| > 1 uses
// pi/2 - atan(1/x).  CPU and GPU produce identical bits.
NF_HD double datan(double x) {
  if (x != x) return x;
  const bool neg = x < 0.0;
  double a = neg ? -x : x;
  const bool inv = a > 1.0;
  if (inv) a = 1.0 / a;
  a = a / (1.0 + sqrt_cr(1.0 + a * a));
  a = a / (1.0 + sqrt_cr(1.0 + a * a));
  const double t = a * a;
  double s = 1.0 / 35.0;
  s = 1.0 / 33.0 - t * s;
  s = 1.0 / 31.0 - t * s;
  s = 1.0 / 29.0 - t * s;
  s = 1.0 / 27.0 - t * s;
  s = 1.0 / 25.0 - t * s;
  s = 1.0 / 23.0 - t * s;
  s = 1.0 / 21.0 - t * s;
  s = 1.0 / 19.0 - t * s;
  s = 1.0 / 17.0 - t * s;
  s = 1.0 / 15.0 - t * s;
  s = 1.0 / 13.0 - t * s;
  s = 1.0 / 11.0 - t * s;
  s = 1.0 / 9.0 - t * s;
  s = 1.0 / 7.0 - t * s;
  s = 1.0 / 5.0 - t * s;
  s = 1.0 / 3.0 - t * s;
  s = 1.0 - t * s;
  double r = 4.0 * (a * s);
  if (inv) r = 1.57079632679489661923 - r;
  return neg ? -r : r;
}

// Van der Waals energy of one contacting pair, the model of
// statKernel_VdW_naive (inc/statKernel_MUT.h:531-628): pairs whose closest
// points are interior to both rods and whose surface gap is below dlim
// contribute  excA / theta / a1 * curve2surface_limit(gap, dlim, a1, a2,
// A, B, theta)  with lengths in Angstrom (x 1e4), theta = atan(a1 / (d - a2
// - dlim)), the excluded area of exclArea_naive (:800-842) and the
// curve-to-surface integral of curve2surface_limit (:891-894).  Rod 2 is
// shifted by s.  Evaluated in FP64 in a fixed order (CPU and GPU agree bit
// for bit).  Deviation: the skew excluded area uses the second rod's radius
// where exclArea_naive reads its length (exclArea_skew, :869-871, uses the
// radius).  Returns false when the pair does not contribute.
NF_HD bool vdw_pair(const float4 P1, const float4 D1, const float4 P2, const float4 D2,
                    double sx, double sy, double sz, double dlim, double A, double B,
                    double* energy) {
  const double PI = 3.14159265358979323846;
  const double c1x = (double)P1.x, c1y = (double)P1.y, c1z = (double)P1.z;
  const double c2x = (double)P2.x + sx, c2y = (double)P2.y + sy, c2z = (double)P2.z + sz;
  const double u1x = (double)D1.x, u1y = (double)D1.y, u1z = (double)D1.z;
  const double u2x = (double)D2.x, u2y = (double)D2.y, u2z = (double)D2.z;
  const double h1 = (double)P1.w, h2 = (double)P2.w, a1 = (double)D1.w, a2 = (double)D2.w;
  const double wx = c1x - c2x, wy = c1y - c2y, wz = c1z - c2z;
  const double aa = u1x * u1x + u1y * u1y + u1z * u1z;
  const double ee = u2x * u2x + u2y * u2y + u2z * u2z;
  const double bb = u1x * u2x + u1y * u2y + u1z * u2z;
  const double cg = bb / sqrt_cr(aa * ee);
  const double sg2 = 1.0 - cg * cg;
  const double sg = sg2 > 0.0 ? sqrt_cr(sg2) : 0.0;
  const double eps = 1e-6;
  double excA = 0.0, dline;
  if (sg > eps) {
    const double d1 = u1x * wx + u1y * wy + u1z * wz;
    const double d2 = u2x * wx + u2y * wy + u2z * wz;
    const double den = aa * ee - bb * bb;
    const double t1 = (bb * d2 - ee * d1) / den;
    const double t2 = (aa * d2 - bb * d1) / den;
    if (!(fabs(t1) <= h1 && fabs(t2) <= h2)) return false;  // closest points not interior
    const double px = (wx + t1 * u1x) - t2 * u2x;
    const double py = (wy + t1 * u1y) - t2 * u2y;
    const double pz = (wz + t1 * u1z) - t2 * u2z;
    dline = sqrt_cr(px * px + py * py + pz * pz);
    const double ea = 2.0 * a1 / sg, eb = 2.0 * a2 / sg;
    excA = ea * eb * sg;
    const double cx = h1 - fabs(t1), cy = h2 - fabs(t2);
    if (fabs(cx) < ea || fabs(cy) < eb) excA = excA / 2.0;
    if (cx <= -ea || cy <= -eb) excA = 0.0;
  } else {
    // parallel: distance of the axes, overlap of the projections on axis 1
    const double n1 = sqrt_cr(aa);
    const double pr = (-(u1x * wx + u1y * wy + u1z * wz)) / n1;  // (c2 - c1) . u1
    const double qx = -wx - pr * u1x / n1, qy = -wy - pr * u1y / n1, qz = -wz - pr * u1z / n1;
    dline = sqrt_cr(qx * qx + qy * qy + qz * qz);
    const double tip0 = fabs(pr) - h2, tip1 = tip0 + 2.0 * h2;
    if (tip1 < -h1 || tip0 > h1) return false;  // axes do not overlap: not interior
    if (tip0 <= -h1 + eps && tip1 >= h1 - eps) excA = 2.0 * a1 * (2.0 * h1);
    if (tip0 > h1 - eps) excA = 0.0;
    if (tip0 > -h1 + eps && tip1 < h1 - eps) excA = 2.0 * 2.0 * a1 * ((tip1 < h1 ? tip1 : h1) - tip0);
  }
  const double gap = (dline - a1) - a2;
  if (!(gap < dlim)) return false;
  const double theta = datan(a1 / ((dline - a2) - dlim));
  // curve2surface_limit in Angstrom
  const double d = gap * 1e4, dl = dlim * 1e4, R1 = a1 * 1e4, R2 = a2 * 1e4;
  const double q = 1.0 - theta * d / dl / dl / (R1 / R2 + 1.0);
  const double d4 = (d * d) * (d * d), dl4 = (dl * dl) * (dl * dl);
  const double d10 = d4 * d4 * (d * d), dl10 = dl4 * dl4 * (dl * dl);
  const double q4 = (q * q) * (q * q);
  const double c2s = -A * PI * R1 / 2.0 * theta * (1.0 / d4 - 1.0 / dl4 * (1.0 / q)) +
                     B * 2.0 * PI * R1 / 5.0 * theta * (1.0 / d10 - 1.0 / dl10 * (1.0 / q4));
  *energy = excA / theta / a1 * c2s;
  return true;
}

// --------------------------------------------------------------------------
// Capsule predicate, FP32 with a rigorous error band.
//
// Returns 1 (d64^2 below the cut: overlap / contact), 0 (not), or -1
// (undecided: the caller recomputes with pair_d2_f64).  `strict` selects
// d^2 < R^2 (hard core) versus d^2 <= R^2 (contact); the result agrees with
// the FP64 rule `within64 && d2_64 (<|<=) R64^2` whenever it is not -1.
// *within is the reach result when decided (meaningful only if ret != -1).
//
// Band.  Let F(s,t) = |w + s h1 - t h2|^2 with exact inputs and F* its
// minimum over the box.  The FP32 point (s,t) is feasible, so
// F* <= F(s,t) <= d2c + E_F, where E_F bounds the evaluation error of v and
// v.v (input rounding of w and h, gamma_3 per dot product).  F is convex, so
// F(s,t) - F* <= 2 sum_i (g_i x_i + |g_i|) (Frank-Wolfe gap over the box,
// g = (h1.v, -h2.v)); the gap is computed in FP32 and inflated by its own
// error bound.  The FP64 value differs from F* by at most EPS64 * S^2
// (S = |w|+|h1|+|h2|), a bound with a margin of >100x over the FP64
// algorithm's worst case.
// --------------------------------------------------------------------------
constexpr float U32 = 5.9604644775390625e-8f;  // 2^-24

// Square root and quotient where only a bound or a search point is needed.
// The device uses the hardware instructions (v_sqrt_f32 and v_rcp_f32,
// each within 1 ulp) instead of the correctly rounded sequences of about a
// dozen instructions; the host uses the library functions.
//   sqrt_up(x) >= sqrt(x) for x >= 0: a root within 1 ulp (2^-23
//   relative) times 1 + 2^-22 (rounding to nearest can only raise the
//   product past that); a zero or denormal x may give 0 on the device,
//   below every margin it is added to.
//   quot(a, b) is a / b within a few ulps: the closest-point parameters of
//   the FP32 band, whose error bounds hold for any parameters in [-1, 1],
//   and the slab parameters of the footprint clips, widened by 8 ulps.
NF_HD float sqrt_up(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x) * (1.0f + 4.0f * U32);
#else
  return sqrtf(x) * (1.0f + 4.0f * U32);
#endif
}

NF_HD float recip(float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(b);
#else
  return 1.0f / b;
#endif
}

NF_HD float quot(float a, float b) { return a * recip(b); }
constexpr float EPS64_REL = 1e-12f;

struct F32Pair {
  int decided;  // -1 undecided, else the cut result (0/1)
  int within;   // reach result when decided
  float d2;     // FP32 squared distance (valid when within)
};

NF_HD F32Pair pair_f32(const float4 P1, const float4 D1, const float4 P2, const float4 D2,
                       const float sx, const float sy, const float sz, const float R,
                       const float ew, const bool strict) {
  F32Pair out;
  const float wx = (P1.x - P2.x) - sx;
  const float wy = (P1.y - P2.y) - sy;
  const float wz = (P1.z - P2.z) - sz;
  const float w2 = fmaf(wz, wz, fmaf(wy, wy, wx * wx));
  const float rr = (P1.w + P2.w) * 1.000001f + R;
  const float rr2 = rr * rr;
  // reach band: w components err <= ew; rr differs from the FP64 reach by
  // < 3.5e-7 relative (5 roundings + the pad constant), so rr^2 by < 8e-7.
  const float nw = sqrt_up(w2) + 2.0f * ew;
  const float ew2 = 1.7321f * ew;
  const float e_w2 = 4.0f * U32 * w2 + (2.0f * nw + ew2) * ew2;
  const float rlo = rr2 * (1.0f - 1.5e-6f), rhi = rr2 * (1.0f + 1.5e-6f);
  if (w2 - e_w2 > rhi) {
    out.decided = 0;
    out.within = 0;
    out.d2 = w2;
    return out;
  }
  if (!(w2 + e_w2 < rlo)) {
    out.decided = -1;
    out.within = 0;
    out.d2 = 0.0f;
    return out;
  }
  out.within = 1;
  const float h1x = P1.w * D1.x, h1y = P1.w * D1.y, h1z = P1.w * D1.z;
  const float h2x = P2.w * D2.x, h2y = P2.w * D2.y, h2z = P2.w * D2.z;
  const float a = fmaf(h1z, h1z, fmaf(h1y, h1y, h1x * h1x));
  const float e = fmaf(h2z, h2z, fmaf(h2y, h2y, h2x * h2x));
  const float b = fmaf(h1z, h2z, fmaf(h1y, h2y, h1x * h2x));
  const float c = fmaf(h1z, wz, fmaf(h1y, wy, h1x * wx));
  const float f = fmaf(h2z, wz, fmaf(h2y, wy, h2x * wx));
  const float denom = fmaf(a, e, -(b * b));
  float s = 0.0f;
  if (denom > 0.0f) {
    s = quot(fmaf(b, f, -(c * e)), denom);
    s = fminf(fmaxf(s, -1.0f), 1.0f);
  }
  float t = quot(fmaf(b, s, f), e);
  if (t < -1.0f) {
    t = -1.0f;
    s = fminf(fmaxf(quot(-b - c, a), -1.0f), 1.0f);
  } else if (t > 1.0f) {
    t = 1.0f;
    s = fminf(fmaxf(quot(b - c, a), -1.0f), 1.0f);
  }
  const float vx = fmaf(-t, h2x, fmaf(s, h1x, wx));
  const float vy = fmaf(-t, h2y, fmaf(s, h1y, wy));
  const float vz = fmaf(-t, h2z, fmaf(s, h1z, wz));
  const float d2 = fmaf(vz, vz, fmaf(vy, vy, vx * vx));
  out.d2 = d2;
  // error bounds
  const float nh1 = P1.w * 1.000001f, nh2 = P2.w * 1.000001f;
  const float S = nw + nh1 + nh2;
  const float Ev = ew2 + U32 * (nh1 + nh2) + 4.1f * U32 * S;
  const float nv = sqrt_up(d2) * 1.000001f + Ev;
  const float EF = 4.0f * U32 * d2 + (2.0f * nv + Ev) * Ev;
  const float gs = fmaf(h1z, vz, fmaf(h1y, vy, h1x * vx));
  const float gt = -fmaf(h2z, vz, fmaf(h2y, vy, h2x * vx));
  const float Egs = 4.1f * U32 * nh1 * nv + 1.01f * nh1 * Ev;
  const float Egt = 4.1f * U32 * nh2 * nv + 1.01f * nh2 * Ev;
  const float ags = fabsf(gs), agt = fabsf(gt);
  const float gap = 2.0f * (fmaf(gs, s, ags) + fmaf(gt, t, agt));
  const float Egap = 4.0f * (Egs + Egt) + 8.0f * U32 * (ags + agt) * 2.0f;
  const float e64 = EPS64_REL * S * S;
  const float band_lo = (EF + gap + Egap + e64) * 1.0001f + 4.0f * U32 * d2;
  const float band_hi = (EF + e64) * 1.0001f + 4.0f * U32 * d2;
  const float R2 = R * R;
  const float R2lo = R2 * (1.0f - 16.0f * U32), R2hi = R2 * (1.0f + 16.0f * U32);
  const float lo = d2 - band_lo, hi = d2 + band_hi;
  if (hi < R2lo) {
    out.decided = 1;
  } else if (strict ? (lo >= R2hi) : (lo > R2hi)) {
    out.decided = 0;
  } else {
    out.decided = -1;
  }
  return out;
}

// --------------------------------------------------------------------------
// Cheap screen in front of pair_f32: the rigorous reach band, then a lower
// bound on the segment distance by the distance of the two infinite lines,
// |w . (u1 x u2)| / |u1 x u2| (Lumelsky 1985: the segments' closest points
// lie on the lines, so the line distance never exceeds the segment
// distance).  Returns
//   SCREEN_UNDECIDED  reach inside its error band: the caller decides reach
//                     and cut in FP64;
//   SCREEN_OUT        not within reach (FP64 agrees);
//   SCREEN_FAR        within reach and certainly not cut: the exact segment
//                     distance exceeds sqrt(R^2 (1 + 16 u) + e64), so the
//                     FP64 squared distance is above R64^2 as well;
//   SCREEN_FULL       within reach, run pair_f32 for the cut.
// Error terms: |w - w_exact| <= sqrt(3) ew; each component of u1 x u2 is
// within 2u of the exact product (|u| <= 1), so |v - v_exact| <= 3.5u; the
// dot product adds 3u |w||v|; e64 = EPS64_REL S^2 bounds the FP64
// algorithm's distance from the exact minimum, as in pair_f32.
// --------------------------------------------------------------------------
enum { SCREEN_UNDECIDED = -1, SCREEN_OUT = 0, SCREEN_FAR = 1, SCREEN_FULL = 2 };

NF_HD int pair_f32_screen(const float4 P1, const float4 D1, const float4 P2, const float4 D2,
                          const float sx, const float sy, const float sz, const float R,
                          const float ew) {
  const float wx = (P1.x - P2.x) - sx;
  const float wy = (P1.y - P2.y) - sy;
  const float wz = (P1.z - P2.z) - sz;
  const float w2 = fmaf(wz, wz, fmaf(wy, wy, wx * wx));
  const float rr = (P1.w + P2.w) * 1.000001f + R;
  const float rr2 = rr * rr;
  const float nw = sqrt_up(w2) + 2.0f * ew;
  const float ew2 = 1.7321f * ew;
  const float e_w2 = 4.0f * U32 * w2 + (2.0f * nw + ew2) * ew2;
  const float rlo = rr2 * (1.0f - 1.5e-6f), rhi = rr2 * (1.0f + 1.5e-6f);
  if (w2 - e_w2 > rhi) return SCREEN_OUT;
  if (!(w2 + e_w2 < rlo)) return SCREEN_UNDECIDED;
  const float vx = fmaf(D1.y, D2.z, -(D1.z * D2.y));
  const float vy = fmaf(D1.z, D2.x, -(D1.x * D2.z));
  const float vz = fmaf(D1.x, D2.y, -(D1.y * D2.x));
  const float v2 = fmaf(vz, vz, fmaf(vy, vy, vx * vx));
  const float t = fmaf(wz, vz, fmaf(wy, vy, wx * vx));
  const float nv = sqrt_up(v2) * (1.0f + 4.0f * U32) + 3.5f * U32;
  const float Et = (3.5f * U32 * nw + ew2 * nv + 3.0f * U32 * nw * nv) * 1.01f;
  const float A = fabsf(t) - Et;
  const float S = nw + P1.w * 1.000001f + P2.w * 1.000001f;
  const float Rb2 = (R * R * (1.0f + 16.0f * U32) + EPS64_REL * S * S) * 1.0001f;
  if (A > 0.0f && A * A > Rb2 * (nv * nv)) return SCREEN_FAR;
  return SCREEN_FULL;
}

// --------------------------------------------------------------------------
// Candidate sampler.  The job of cuda_CNTgenerate_orient
// (inc/cuda_CNTgenerate.h:136-260): uniform position, von Neumann accept
// against the density target (:227-232), orientation from the target law
// (genCosine/inc_rotation :262-295, here Marsaglia 1972 for isotropic and
// Wood 1994 for vMF), length from the length law (reference: constant
// DEF_L, :246).  Returns 1 proposal, 0 density reject, -1 sampler exhausted,
// -2 overlaps its own periodic image.
// --------------------------------------------------------------------------
NF_HD bool self_image_overlap(const SampleDesc& d, const float4 P, const float4 D) {
  const double R = d.soft_frac * ((double)D.w + (double)D.w) + d.sep;
  const double rr = ((double)P.w + (double)P.w) * REACH_PAD + R;
  double lmin = d.box[0];
  if (d.box[1] < lmin) lmin = d.box[1];
  if (d.box[2] < lmin) lmin = d.box[2];
  if (rr < lmin) return false;
  int K[3];
  for (int a = 0; a < 3; ++a) K[a] = (int)floor(rr / d.box[a]) + 1;
  for (int nz = -K[2]; nz <= K[2]; ++nz)
    for (int ny = -K[1]; ny <= K[1]; ++ny)
      for (int nx = -K[0]; nx <= K[0]; ++nx) {
        if (nx == 0 && ny == 0 && nz == 0) continue;
        bool within;
        const double d2 = pair_d2_f64(P, D, P, D, (double)nx * d.box[0], (double)ny * d.box[1],
                                      (double)nz * d.box[2], R, &within);
        if (within && d2 < R * R) return true;
      }
  return false;
}

NF_HD int sample_candidate(const SampleDesc& d, uint32_t round, uint32_t slot, Rod* out) {
  const double Lx = d.box[0], Ly = d.box[1], Lz = d.box[2];
  u32x4 o = philox(slot, round, 0u, 0u, d.seed, d.stream);
  const double ux = u01(o.x, o.y), uy = u01(o.z, o.w);
  o = philox(slot, round, 1u, 0u, d.seed, d.stream);
  const double uz = u01(o.x, o.y), ud = u01(o.z, o.w);
  float xf = (float)(ux * Lx), yf = (float)(uy * Ly), zf = (float)(uz * Lz);
  if (!(xf < (float)Lx)) xf = 0.0f;
  if (!(yf < (float)Ly)) yf = 0.0f;
  if (!(zf < (float)Lz)) zf = 0.0f;

  if (d.den_law == NF_DEN_AGGLOM) {
    const double x = (double)xf, y = (double)yf, z = (double)zf;
    double d2min = nf_bits_to_double(0x7FF0000000000000ull);
    for (int k = 0; k < d.n_agg; ++k) {
      double dx = x - d.agg[3 * k + 0], dy = y - d.agg[3 * k + 1], dz = z - d.agg[3 * k + 2];
      if (dx > 0.5 * Lx) dx -= Lx; else if (dx < -0.5 * Lx) dx += Lx;
      if (dy > 0.5 * Ly) dy -= Ly; else if (dy < -0.5 * Ly) dy += Ly;
      if (dz > 0.5 * Lz) dz -= Lz; else if (dz < -0.5 * Lz) dz += Lz;
      const double dd = (dx * dx + dy * dy) + dz * dz;
      if (dd < d2min) d2min = dd;
    }
    const double inv2s2 = 1.0 / (2.0 * d.agg_sigma * d.agg_sigma);
    const double rho = 1.0 + (d.agg_peak - 1.0) * dexp(-(d2min * inv2s2));
    if (!(ud * d.agg_peak < rho)) return 0;
  } else if (d.den_law == NF_DEN_MESH) {
    // the map's values are at most 1: accept iff u < rho(x) (cuda_CNTgenerate.h:227-232)
    const double rho = mesh_density(d.mesh, d.mext, d.box, (double)xf, (double)yf, (double)zf);
    if (!(ud < rho)) return 0;
  }

  double cx = 0.0, cy = 0.0, cz = 0.0;
  bool ok = false;
  if (d.ort_law == NF_ORT_VMF) {
    o = philox(slot, round, 8u, 0u, d.seed, d.stream);
    const double uw = u01(o.x, o.y);
    const double em2k = dexp(-2.0 * d.kappa);
    double w = 1.0 + dlog(uw + (1.0 - uw) * em2k) / d.kappa;
    if (w > 1.0) w = 1.0;
    if (w < -1.0) w = -1.0;
    for (uint32_t t = 0; t < 32 && !ok; ++t) {
      o = philox(slot, round, 9u + t, 0u, d.seed, d.stream);
      const double v1 = 2.0 * u01(o.x, o.y) - 1.0, v2 = 2.0 * u01(o.z, o.w) - 1.0;
      const double s = v1 * v1 + v2 * v2;
      if (s < 1.0 && s > 0.0) {
        const double rr = sqrt_cr(1.0 - w * w) / sqrt_cr(s);
        const double a1 = v1 * rr, a2 = v2 * rr;
        cx = (w * d.axis[0] + a1 * d.e1[0]) + a2 * d.e2[0];
        cy = (w * d.axis[1] + a1 * d.e1[1]) + a2 * d.e2[1];
        cz = (w * d.axis[2] + a1 * d.e1[2]) + a2 * d.e2[2];
        ok = true;
      }
    }
  } else {
    for (uint32_t t = 0; t < 32 && !ok; ++t) {
      o = philox(slot, round, 8u + t, 0u, d.seed, d.stream);
      const double v1 = 2.0 * u01(o.x, o.y) - 1.0, v2 = 2.0 * u01(o.z, o.w) - 1.0;
      const double s = v1 * v1 + v2 * v2;
      if (s < 1.0 && s > 0.0) {
        const double q = 2.0 * sqrt_cr(1.0 - s);
        cx = v1 * q;
        cy = v2 * q;
        cz = 1.0 - 2.0 * s;
        ok = true;
      }
    }
  }
  if (!ok) return -1;

  double L = d.len;
  if (d.len_law == NF_LEN_LOGNORMAL) {
    ok = false;
    double z = 0.0;
    for (uint32_t t = 0; t < 32 && !ok; ++t) {
      o = philox(slot, round, 48u + t, 0u, d.seed, d.stream);
      const double v1 = 2.0 * u01(o.x, o.y) - 1.0, v2 = 2.0 * u01(o.z, o.w) - 1.0;
      const double s = v1 * v1 + v2 * v2;
      if (s < 1.0 && s > 0.0) {
        z = v1 * sqrt_cr(-2.0 * dlog(s) / s);
        ok = true;
      }
    }
    if (!ok) return -1;
    L = d.len * dexp(d.len_sigma * z);
    if (L < d.len_min) L = d.len_min;
    if (L > d.len_max) L = d.len_max;
  }
  const float lf = (float)L;
  out->P = make_float4(xf, yf, zf, 0.5f * lf);
  out->D = make_float4((float)cx, (float)cy, (float)cz, (float)d.radius);
  if (self_image_overlap(d, out->P, out->D)) return -2;
  return 1;
}

}  // namespace nf
