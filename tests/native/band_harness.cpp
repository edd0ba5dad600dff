// Host build of the library's pair predicates (test infrastructure): lets
// the CPU test suite fuzz the FP32 error band against the FP64 oracle.
#include "../../nanofiller_amd/csrc/nf_common.h"

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

}  // extern "C"
