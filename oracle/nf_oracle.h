/*
 * nf_oracle.h - FP64 CPU oracle of the NanoFiller hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library calls, links or
 * loads this code; it is the checker that the HIP kernels are compared
 * against (tests/, the smoke check and the CPU-baseline leg of bench.py).
 *
 * It restates, in plain single-threaded (optionally OpenMP) C, the
 * generation, contact, histogram and network algorithms of the MI355X
 * library at the level of individual IEEE operations, so that every
 * accept/reject decision, every contact pair and every integer count can be
 * compared bit-for-bit.  Each function cites the reference routine
 * (pyrlinum/NanoFiller, file:line) whose behaviour it follows, and the
 * documented deviations from it (exact segment-segment distance instead of
 * the single-pass clamp, priority round rule instead of launch-order
 * dependent removal, Philox instead of clock()-seeded XORWOW).
 */
#ifndef NF_ORACLE_H
#define NF_ORACLE_H

#include <stdint.h>

#include "../include/nanofiller.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 (Salmon et al., SC'11): out = philox(ctr, key). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 52-bit uniform double in (0,1) from two 32-bit words. */
double orc_u01(uint32_t a, uint32_t b);
/* Deterministic exp / log built from +, -, *, / only. */
double orc_exp(double x);
double orc_log(double x);

/* Agglomerate centres (3 doubles each) of instance p. */
void orc_agg_centres(const nf_params* p, double* out);
double orc_mesh_density(const nf_params* p, double x, double y, double z);
void orc_node_points(const nf_params* p, const float* rods, uint32_t n, const uint32_t* I,
                     const uint32_t* J, const int32_t* img, uint64_t nc, double* tI, double* tJ,
                     uint32_t* flag, int threads);
/* Orthonormal basis e1, e2 perpendicular to the (unit) vMF axis. */
void orc_vmf_basis(const double mu[3], double e1[3], double e2[3]);

/* Candidate of slot `slot` in round `round`: 1 = proposal (rod in out),
 * 0 = rejected by the density test, -1 = sampler exhausted, -2 = overlaps
 * its own periodic image. */
int orc_candidate(const nf_params* p, const double* agg, uint32_t round, uint32_t slot,
                  float out[8]);

/* Squared distance between the axis segments of capsules ri and rj, where
 * rj is displaced by shift (its periodic image).  *within receives the
 * centre-distance reach test for hard-core/contact distance R. */
double orc_pair_d2(const float* ri, const float* rj, const double shift[3], double R,
                   int* within);

/* Contact area S of capsules ri and rj (rj displaced by shift) with contact
 * gap rlim, and of a contact list (I, J, image as in orc_contacts). */
double orc_contact_area(const float* ri, const float* rj, const double shift[3], double rlim);
void orc_contact_areas(const nf_params* p, const float* rods, const uint32_t* I, const uint32_t* J,
                       const int32_t* image, int64_t nc, double rlim, double* out);

/* Batch of independent pair evaluations (tests). */
void orc_pair_batch(int n, const float* ra, const float* rb, const double* sh, int contact,
                    double sep_or_gap, double soft_frac, double* d2out, int* within_out,
                    int* cut_out);

typedef struct orc_gen_stats {
  uint32_t n_accepted, rounds;
  int32_t status;
  uint32_t pad;
  uint64_t candidates, proposals, pair_tests, rejected_old, rejected_new;
} orc_gen_stats;

/* Full insert/reject generation.  rods receives n_rods*8 floats in id order.
 * round_accepts (may be NULL, max_rounds entries) receives the per-round
 * accepted counts.  threads <= 1: single-threaded. */
int orc_generate(const nf_params* p, float* rods, orc_gen_stats* st,
                 uint32_t* round_accepts, int threads);

/* Contacts with surface gap <= gap (I < J, all periodic images), sorted by
 * (I, J, image).  Returns the count; arrays are filled up to cap entries. */
int64_t orc_contacts(const nf_params* p, const float* rods, uint32_t n, double gap,
                     uint32_t* I, uint32_t* J, float* gap32, double* gap64,
                     int32_t* image, int64_t cap, uint64_t* pair_tests, int threads);

/* Histograms (exact int64): centre along axis, polar angle from pol_axis,
 * mutual angle of the given contact pairs, phi-theta 2-D (optional). */
int orc_histograms(const nf_params* p, const float* rods, uint32_t n,
                   const uint32_t* I, const uint32_t* J, int64_t nc,
                   const nf_hist_spec* spec, int64_t* centre, int64_t* polar,
                   int64_t* mutual, int64_t* phitheta);

/* Network: component labels (smallest id), CSR Kirchhoff matrix on N+2
 * nodes.  Returns nnz (or <0 on error); arrays filled when non-NULL and
 * cap is large enough. */
int64_t orc_network(const nf_params* p, const float* rods, uint32_t n,
                    const uint32_t* I, const uint32_t* J, const double* gap64,
                    const int32_t* image, int64_t nc, uint32_t* labels,
                    uint32_t* row_ptr, uint32_t* col, double* val, int64_t cap,
                    nf_network_info* info);

/* Dead-end pruning (reduce): keep[i] = 1 for the kept rods; returns their
 * number, *rounds the iterations including the last, empty one. */
/* expand (simPrms::expand, src/simPrms.cpp:744-852, its copies chosen by
 * centre): periodic copies along axis k of the rods whose centre x_k lies
 * within B of the far face (x_k >= L - B, copied below the box) and of the
 * near face (x_k < B, copied above it); every coordinate along k shifted by
 * B, each as one FP64 sum rounded to FP32 (x + B, x + (B - L), x + (B + L)).
 * out: the n rods in id order, then the copies below, then those above,
 * each in id order.  Returns the rod count; writes only when cap suffices. */
int64_t orc_expand(const float* rods, uint32_t n, int k, double L, double B, float* out, int64_t cap);
uint32_t orc_reduce(const nf_params* p, const float* rods, uint32_t n, const uint32_t* I,
                    const uint32_t* J, const int32_t* image, int64_t nc, double gcut,
                    int32_t min_cont, uint8_t* keep, uint32_t* rounds);

/* Deterministic atan (+, -, *, / and the correctly rounded sqrt). */
double orc_atan(double x);
/* Van der Waals energy of one pair / summed over a sorted contact list. */
int orc_vdw_pair(const float* r1, const float* r2, const double sh[3], double dlim, double A,
                 double B, double* energy);
uint64_t orc_vdw(const nf_params* p, const float* rods, uint32_t n, const uint32_t* I,
                 const uint32_t* J, const int32_t* image, int64_t nc, double dlim, double A,
                 double B, double* energy);
/* Centre-centre distance histogram and exact fixed-point sums (brute force). */
uint64_t orc_centre_distance(const nf_params* p, const float* rods, uint32_t n, int32_t bins,
                             double width, int64_t* counts, uint64_t* sum_q, uint64_t* sum2_q);

#ifdef __cplusplus
}
#endif
#endif
