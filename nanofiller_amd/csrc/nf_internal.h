// nf_internal.h - device descriptors, per-instance state and launcher
// declarations shared by the HIP translation units and the host library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "nf_common.h"
#include "nf_seg.h"

namespace nf {

constexpr int WAVE = 64;
constexpr float CELL_MARGIN = 1e-4f;  // rods may sit this far outside their cell

// Per-instance constant descriptor (device array, read by every kernel).
struct InstDesc {
  SampleDesc smp;
  float boxf[3];
  int32_t ncell[3];
  double cell_scale[3];  // ncell / box, for cell assignment
  float edge[3];         // cell edge
  uint32_t cell_base;    // first global cell of this instance
  uint32_t ncells;
  uint32_t cap;          // accepted-rod capacity per cell
  uint32_t rod_off;      // offset of this instance in the by-id rod arrays
  uint64_t old_off;      // offset of this instance in the cell-slot arrays
  uint32_t n_target, min_batch, max_batch, max_rounds, stall_rounds;
  uint32_t st_off, st_n;    // generation stencil (int4 entries), full
  uint32_t hst_off, hst_n;  // half stencil of the same-round pass (each cell pair once)
  double sf64, sep64;     // hard core = sf*(a1+a2)+sep
  float sf32, sep32;
  float ew;               // FP32 bound on a centre-difference component error
  float hl_max;           // max half length of any rod of this instance
  float R_max;            // max hard-core distance
  uint32_t pad;
};

// Per-instance dynamic state (device), updated by the round controller.
struct InstState {
  uint32_t n_acc, round, M, status;
  uint32_t mprev, aprev, stall, overflow;
  unsigned long long candidates, proposals, pair_tests, rechecks, rej_old, rej_new;
  double max_dd;  // diagnostic max |d32 - d64|
  unsigned long long tile_bytes;      // probe + tile bytes streamed by the reject-vs-accepted kernel
  unsigned long long tile_bytes_new;  // same, reject-vs-same-round kernel
};

struct RoundInfo {
  uint32_t T;         // candidate slots this round over all instances
  uint32_t running;   // instances still running after the last round
  uint32_t rounds;    // rounds executed by the controller
  uint32_t est;       // rounds still needed if the last round's acceptance holds (max over instances)
  uint32_t old;       // instances with candidates and accepted rods this round (0: no first-pass pairs)
  uint32_t pad[3];
};

// Wave reductions for the end of a kernel (every lane converged): sum and
// max over the 64 lanes, result in every lane.
// Cell of a centre in a cell list binned with nc cells and scale = nc / box
// per axis (FP64 product, truncated, clamped to the grid): the binning and
// every pass that needs a listed rod's cell use this one function.
__device__ __forceinline__ void cell_coords(const float4 P, const int32_t nc[3], const double sc[3], int c[3]) {
  const float p[3] = {P.x, P.y, P.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int i = (int)((double)p[a] * sc[a]);
    c[a] = i < 0 ? 0 : (i >= nc[a] ? nc[a] - 1 : i);
  }
}

__device__ __forceinline__ uint32_t cell_index(const float4 P, const int32_t nc[3], const double sc[3]) {
  int c[3];
  cell_coords(P, nc, sc, c);
  return (uint32_t)(c[0] + nc[0] * (c[1] + nc[1] * c[2]));
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, WAVE);
    v = w > v ? w : v;
  }
  return v;
}

// Work queues of the pair kernels: 64 per launch, each counter on its own
// 128-byte line (QSTRIDE words apart) so that grabs on different queues do
// not serialise on one line.  Queue q covers the cell range q of 64 equal
// ranges; the eight queues of XCD x are the ranges 8x .. 8x+7.
constexpr int QUEUES_PER_KERNEL = 64;
constexpr int QSTRIDE = 32;
constexpr int QWORDS = QUEUES_PER_KERNEL * QSTRIDE;  // words per launch
enum { Q_OLD = 0, Q_NEW = 1, Q_CNT = 2, Q_REC = 3, Q_N = 4 };

// Device buffers of a generation batch.
// Launch shape of the generation's pair passes (0 / -1: the built-in
// choice of pair_grid, pair_part_shift and pair_split_min).
struct PairTiles {
  uint32_t waves_old = 0;  // waves per CU, reject vs accepted rods
  uint32_t waves_new = 0;  // waves per CU, reject vs same-round proposals
  int32_t part_shift = -1; // work items per dense cell (reject vs accepted), as a shift
  uint32_t split_min = 0;  // probes per part of a split cell, at least
  uint32_t test_all = 0;   // reject vs accepted: cut proposals stay in the walk (every pair within reach tested)
};

struct GenBuffers {
  PairTiles tiles;
  InstDesc* desc = nullptr;
  InstState* st = nullptr;
  RoundInfo* ri = nullptr;
  uint32_t* slot_base = nullptr;  // n_inst+1, exclusive prefix of M
  int4* stencil = nullptr;
  double* agg = nullptr;
  float* mesh = nullptr;          // density maps of the NF_DEN_MESH instances
  uint32_t* cell_inst = nullptr;  // instance of every global cell
  // accepted rods in fixed-capacity cell slots
  float4* oP = nullptr;
  float4* oD = nullptr;
  uint32_t* oid = nullptr;
  uint32_t* ostart = nullptr;  // start slot of every cell
  uint32_t* ocount = nullptr;
  float* ohl = nullptr;        // max half length in the cell
  // accepted rods by id
  float4* bP = nullptr;
  float4* bD = nullptr;
  // per-round candidate staging, indexed by global slot
  float4* sP = nullptr;
  float4* sD = nullptr;
  int32_t* scell = nullptr;
  uint32_t* srk = nullptr;
  uint32_t* surv = nullptr;   // T_cap + 1
  uint32_t* srank = nullptr;  // T_cap + 1
  // proposals sorted by cell
  float4* pP = nullptr;
  float4* pD = nullptr;
  uint32_t* pslot = nullptr;
  uint32_t* palive = nullptr;
  uint32_t* pcount = nullptr;  // total_cells
  uint32_t* pstart = nullptr;  // total_cells + 1
  uint32_t* order = nullptr;   // total_cells: work order of a pair pass
  // survivors of the reject-vs-accepted pass, compacted per cell
  float4* aP = nullptr;
  float4* aD = nullptr;
  uint32_t* aslot = nullptr;   // T_cap
  uint32_t* aidx = nullptr;    // T_cap + 1: exclusive prefix of the alive flags
  uint32_t* acount = nullptr;  // total_cells
  uint32_t* astart = nullptr;  // total_cells + 1
  float* phl = nullptr;
  uint32_t* queues = nullptr;  // Q_N * QWORDS
  uint32_t* scan_tmp = nullptr;
  uint64_t total_cells = 0, total_old = 0, T_cap = 0, total_rods = 0;
  int n_inst = 0;
};

// Contact / network buffers of a single packing.
struct ContactBuffers {
  // compact cell list
  float4* cP = nullptr;
  float4* cD = nullptr;
  uint32_t* cid = nullptr;
  uint32_t* cstart = nullptr;  // ncells + 1
  uint32_t* ccount = nullptr;
  float* chl = nullptr;
  uint32_t* cmask = nullptr;  // footprint engine: near mask of each rod of the list
  double cscale[3] = {0.0, 0.0, 0.0};  // cells per unit length (the list's binning)
  int4* stencil = nullptr;
  uint32_t st_n = 0;
  int32_t ncell[3] = {0, 0, 0};
  uint32_t ncells = 0;
  // contacts
  uint32_t* cnt = nullptr;     // N
  uint32_t* rowptr = nullptr;  // N + 1
  uint32_t* fill = nullptr;    // N
  uint32_t* J = nullptr;
  float* gap = nullptr;
  double* gap64 = nullptr;
  int32_t* img = nullptr;
  double* area = nullptr;      // contact (tunnelling) area S of each contact
  uint64_t n_contacts = 0, cap_contacts = 0;
  uint32_t n_rods = 0;
  double gap_cut = 0.0;
  float hl_max = 0.0f, a_max = 0.0f;  // bounds on half length and radius of the rods
  int engine = 0;                     // NF_ENGINE_* that found the last contact list
  int cell_list = 0;                  // the centre cell list held: 1 cell-pair engine's, 2 footprint engine's
  double fp_edge = 0.0;               // footprint engine: home cell edge (um) and fine cells
  int fp_fine = 0;                    // per axis of the last pass (0: cell engine)
  bool valid = false;
  // network
  uint32_t* labels = nullptr;
  uint32_t* nrow = nullptr;
  uint32_t* ncol = nullptr;
  double* nval = nullptr;
  uint64_t nnz = 0;
  bool net_valid = false;
  // node-level segment network (nf_node_network)
  uint32_t *nd_pflag = nullptr, *nd_edge_a = nullptr, *nd_edge_b = nullptr;
  uint32_t *nd_esrc = nullptr, *nd_esnk = nullptr, *nd_src_node = nullptr, *nd_snk_node = nullptr;
  double *nd_egs = nullptr, *nd_egk = nullptr;
  uint32_t* nd_node_rod = nullptr;
  double* nd_node_t = nullptr;
  uint32_t *nd_seg_a = nullptr, *nd_seg_b = nullptr;
  double *nd_seg_L = nullptr, *nd_seg_A = nullptr;
  uint32_t *nd_ce_a = nullptr, *nd_ce_b = nullptr, *nd_ce_k = nullptr;  // kept edges, compacted
  uint32_t* nd_cel_node = nullptr;                                      // electrode points, compacted
  double* nd_cel_gap = nullptr;
  int32_t* nd_cel_side = nullptr;
  uint64_t nd_n_edges = 0, nd_n_elec = 0;
  uint64_t nd_nodes = 0, nd_segments = 0, nd_points = 0;
  bool nd_valid = false;
};

// Arguments of the cell-pair engine.
struct PairArgs {
  const InstDesc* desc;
  const RoundInfo* ri;  // generation: this round's plan (T == 0: no work)
  InstState* st;
  const uint32_t* cell_inst;
  uint32_t ncells_total;
  uint32_t part_shift;  // work items per cell: 1 << part_shift (pair_part_shift)
  uint32_t split_min;   // probes per part of a split cell, at least (pair_split_min)
  uint32_t test_all;    // reject vs accepted: test every pair within reach (NF_OPT_TEST_ALL_PAIRS)
  uint32_t* order;      // ncells_total: work order of the cells (filled by launch_pair; nullptr: none)
  const int4* stencil;
  uint32_t st_off, st_n;  // used when desc == nullptr (contacts)
  // probes
  const float4* prP;
  const float4* prD;
  const uint32_t* prX;
  const uint32_t* prStart;
  const uint32_t* prCount;
  const float* prHl;  // longest probe per cell (reject-vs-accepted pass run accepted-major)
  // neighbours
  const float4* nbP;
  const float4* nbD;
  const uint32_t* nbX;
  const uint32_t* nbStart;
  const uint32_t* nbCount;
  const float* nbHl;
  uint32_t* qctr;
  // outputs
  uint32_t* alive_out;
  uint32_t* surv;
  // contacts: unsorted append buffer (capacity app_cap, count *app_n)
  uint32_t* app_n;
  uint32_t app_cap;
  uint32_t* outI;
  uint32_t* outJ;
  float* outGap;
  double* outGap64;
  int32_t* outImg;
  // contact geometry (single instance)
  double box[3];
  float boxf[3];
  int32_t ncell[3];
  double gap64;
  float gap32;
  float ew;
  float hl_max, R_max;
  unsigned long long* ctr;  // [0] pair tests, [1] rechecks, [2] max dd bits, [3] tile bytes
  int diag;                 // record max |d32-d64| over full tests
};

enum PairMode { PM_GEN_OLD = 0, PM_GEN_NEW = 1, PM_CNT = 2 };

// Arguments of the footprint contact pass (nf_seg.hip).
// Home cells of one work unit of the footprint pass: a run along x.
constexpr int SEG_ROW = 4;

struct SegCrossArgs {
  // centre cell list of the contact rods
  const float4* cP;
  const float4* cD;
  const uint32_t* cid;
  const uint32_t* cstart;
  const uint32_t* ccount;
  const float* chl;
  const uint32_t* cmask;  // near mask of each rod of the list (seg_near_mask)
  const int4* stencil;  // stencil of a SEG_ROW-cell home run, sorted by minimum distance
  uint32_t st_n;
  float reach2;  // (hl_max + rpad)^2: the stencil walk stops past it
  float rpad;    // footprint radius plus the cell margin of a centre
  int fine;      // fine cells per axis of a home cell (1..4)
  SegGrid g;     // the cell grid (claim rule) and footprint radius
  SegCtx x;
  uint32_t* next;      // eight work counters over the work units, QSTRIDE words apart
  uint32_t* overflow;  // set when a home cell has more survivors than fit
  uint32_t* app_n;
  uint32_t app_cap;
  uint32_t* outI;
  uint32_t* outJ;
  float* outGap;
  double* outGap64;
  int32_t* outImg;
  unsigned long long* ctr;  // [0] capsule tests, [1] rechecks, [2] max dd bits, [3] bytes,
                            // [5] units redone one cell at a time
};

// ---- launchers (nf_scan.hip) ----
void scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s);
size_t scan_tmp_words(uint64_t n);
void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);

// ---- launchers (nf_gen.hip) ----
void launch_round(GenBuffers& b, int prec, int diag, hipStream_t s, hipEvent_t* ev);
void launch_ctl_init(GenBuffers& b, hipStream_t s);

// ---- launchers (nf_pair.hip) ----
void launch_pair(int mode, int prec, const PairArgs& a, hipStream_t s, const PairTiles* tiles = nullptr);
// Persistent grid of a pair pass (waves): 16 per CU, the most that the
// register budget lets reside (measured on MI355X, 500k-rod workload in 3
// rounds: same-round pass 1.72 ms per generation at 16 per CU, 1.82 at 12;
// with the earlier, smaller rounds 12 per CU had been best).
int pair_grid(int mode);
// Work items per cell of a pair pass, as a shift: each dense cell's stencil
// walk is dealt over 1 << shift items, so that dense cells do not hold a
// launch back.  launch_pair applies it to launches with fewer than 16 cells
// per wave of the grid.
int pair_part_shift(int mode);
int pair_split_min(int mode);

// ---- launchers (nf_seg.hip) ----
void launch_seg_cross(int prec, const SegCrossArgs& a, hipStream_t s);
// Near masks (seg_near_mask) of the n rods of a cell list binned with
// (nc, scale); h and r are the footprint grid's cell edges and radius.
void launch_seg_marks(const float4* cP, const float4* cD, uint32_t n, const int32_t nc[3], const double scale[3],
                      const float h[3], float r, float mx, uint32_t* cmask, hipStream_t s);

// ---- launchers (nf_post.hip) ----
// Cell list of n rods (two-level counting sort, nf_post.hip): bin_plan
// sizes the scratch (M: m_words + 1 words, scan_tmp: scan_tmp_words(m_words),
// tP/tD/tid: n each), launch_bin_rods writes the rods in cell order (oP, oD,
// oid), each cell's start (ncells + 1 entries), count and largest half
// length.
struct BinPlan {
  uint32_t shift;  // log2 of the cells per bucket
  uint32_t nb;     // buckets
  uint32_t nblk;   // blocks of the histogram and scatter passes
  uint32_t chunk;  // rods per block
  uint64_t m_words;
};
BinPlan bin_plan(uint32_t n, uint64_t ncells);
void launch_bin_rods(const BinPlan& bp, const float4* P, const float4* D, uint32_t n, const int32_t ncell[3],
                     const double scale[3], uint32_t* M, uint32_t* scan_tmp, float4* tP, float4* tD,
                     uint32_t* tid, float4* oP, float4* oD, uint32_t* oid, uint32_t* start, uint32_t* count,
                     float* hl, hipStream_t s);
void launch_expand_flags(const float4* P, uint32_t n, int k, double L, double B, uint32_t* below, uint32_t* above,
                         hipStream_t s);
void launch_expand_scatter(const float4* P, const float4* D, uint32_t n, int k, double L, double B,
                           const uint32_t* below, const uint32_t* rank_below, const uint32_t* above,
                           const uint32_t* rank_above, uint32_t n_below, float4* oP, float4* oD, hipStream_t s);
void launch_row_sort(const uint32_t* rowptr, uint32_t n, uint32_t* J, float* gap, double* gap64,
                     int32_t* img, hipStream_t s);
// Contact area S (contact_area_f64) of every contact of the CSR, rod J at
// its image; rlim is the contact gap.
void launch_contact_area(const float4* P, const float4* D, uint32_t n, const uint32_t* rowptr,
                         const uint32_t* J, const int32_t* img, const double box[3], double rlim,
                         double* area, hipStream_t s);
// CSR over I from the unsorted contact list: per-row counts, then scatter
// (rowptr from scan_u32 of the counts in between).
void launch_csr_count(const uint32_t* I, uint64_t c, uint32_t* cnt, hipStream_t s);
void launch_csr_scatter(const uint32_t* I, const uint32_t* J, const float* gap, const double* gap64,
                        const int32_t* img, uint64_t c, const uint32_t* rowptr, uint32_t* fill,
                        uint32_t* oJ, float* ogap, double* ogap64, int32_t* oimg, hipStream_t s);
void launch_hist(const float4* P, const float4* D, uint32_t n, const uint32_t* rowptr,
                 const uint32_t* J, const double* edges_p, const double* edges_m,
                 const double* edges_t, int nb, int npb, int ntb, int axis, double L,
                 const double* mu, unsigned long long* out, hipStream_t s);
void launch_selftest(int which, const double* in, double* out, uint64_t n, hipStream_t s);
void launch_selftest_pairs(const float* ra, const float* rb, const double* sh, uint64_t n, int contact,
                           double sep_or_gap, double soft_frac, float ew, int32_t* out_screen, int32_t* out_dec,
                           int32_t* out_within, hipStream_t s);
constexpr int CENT_BINS_MAX = 1024;
void launch_cent(const float4* P, const uint32_t* X, const uint32_t* start, const uint32_t* count,
                 const int4* st, uint32_t st_n, const int32_t nc[3], const double box[3],
                 const double* e2, int bins, float r2f, unsigned long long* out, hipStream_t s);
void launch_vdw(const float4* P, const float4* D, uint32_t n, const uint32_t* rowptr,
                const uint32_t* J, const int32_t* img, const double box[3], double dlim, double A,
                double B, double* rowsum, double* bsum, double* out, unsigned long long* pairs,
                hipStream_t s);

// ---- conductance solve (nf_solve.hip) ----
constexpr int CG_MAX_BLOCKS = 1024;
// Device scalars of the conjugate-gradient solve.
struct CgScal {
  double rz, pq, alpha, beta, rr, bb;  // r.z, p.Ap, step, direction update, r.r, b.b
  double isrc, isnk;                   // currents into the source and out to the sink
};
struct CgArgs {
  uint32_t m;  // unknowns
  const uint32_t* crow;
  const uint32_t* ccol;
  const double* cval;
  const double* dinv;
  const double* bsrc;  // source couplings (right-hand side)
  const double* bsnk;  // sink couplings
  double *x, *r, *z, *p, *q;
  double* part;  // 2 * CG_MAX_BLOCKS block partials
  CgScal* scal;
};
int cg_grid(uint64_t m);
void launch_cs_flag(const uint32_t* nrow, const uint32_t* ncol, const uint32_t* labels, uint32_t n,
                    uint32_t* flag, hipStream_t s);
void launch_cs_active(const uint32_t* nrow, const uint32_t* ncol, const uint32_t* labels,
                      const uint32_t* flag, uint32_t n, uint32_t* act, uint32_t* rl, hipStream_t s);
void launch_cs_fill(const uint32_t* nrow, const uint32_t* ncol, const double* nval, uint32_t n,
                    const uint32_t* act, const uint32_t* idx, const uint32_t* off, uint32_t* crow,
                    uint32_t* ccol, double* cval, double* dinv, double* bsrc, double* bsnk,
                    hipStream_t s);
void launch_cs_potential(uint32_t n, const uint32_t* labels, const uint32_t* flag,
                         const uint32_t* act, const uint32_t* idx, const double* x, double* v,
                         hipStream_t s);
// Unknowns in the order of ord (n rods): stage 0 gathers act and rl into
// acto and rlo; stage 1 scatters their scans (n + 1 entries) back to idx and
// off per rod, the totals at n.
void launch_cs_order(uint32_t n, const uint32_t* ord, const uint32_t* act, const uint32_t* rl, uint32_t* acto,
                     uint32_t* rlo, const uint32_t* idxo, const uint32_t* offo, uint32_t* idx, uint32_t* off,
                     int stage, hipStream_t s);
void launch_cg_init(const CgArgs& a, hipStream_t s);
void launch_cg_iteration(const CgArgs& a, hipStream_t s);
void launch_cg_current(const CgArgs& a, hipStream_t s);

// ---- node-level segment network (nf_nodes.hip) ----
enum NodeFlag { NODE_KEPT = 0, NODE_SCREENED = 1, NODE_CROSSING = 2 };
enum PointKind { PT_CONTACT_I = 0, PT_CONTACT_J = 1, PT_SOURCE = 2, PT_SINK = 3 };
struct NodeArgs {
  uint32_t n;
  // contacts: CSR over the lower id, J > row, image applies to J
  const uint32_t* rowptr;
  const uint32_t* rowof;  // row of every contact (filled by launch_node_points)
  const uint32_t* J;
  const int32_t* img;
  const float4* P;  // rods by id
  const float4* D;
  int axis;
  double box[3];
  double L, gcut, hl_max, a_max;
  // contact cell grid (rods sorted by cell)
  const float4* cP;
  const float4* cD;
  const uint32_t* cid;
  const uint32_t* cstart;
  const uint32_t* ccount;
  int32_t ncell[3];
  double cell_scale[3];
  // per contact
  uint32_t* pflag;
  double *ptI, *ptJ;
  uint32_t *edge_a, *edge_b;
  // per rod
  uint32_t *npt, *pstart, *pfill, *ngrp, *nseg, *gstart, *sstart;
  uint32_t *esrc, *esnk, *src_node, *snk_node;
  double *egs, *egk;
  // points
  double* pt_t;
  unsigned long long* pt_key;  // kind << 32 | contact or rod index
  uint32_t* pt_grp;
  // nodes and segments
  uint32_t* node_rod;
  double* node_t;
  uint32_t *seg_a, *seg_b;
  double *seg_L, *seg_A;
};
// Contact points and screening; rowof (one entry per contact) is filled first.
void launch_node_points(const NodeArgs& a, uint32_t* rowof, hipStream_t s);
void launch_node_lists(const NodeArgs& a, hipStream_t s);
void launch_node_nodes(const NodeArgs& a, hipStream_t s);
// out[4] (zeroed by the caller): kept contacts, screened contacts, source and sink points.
void launch_node_kept_flag(const uint32_t* pflag, uint32_t nc, uint32_t* f, hipStream_t s);
void launch_node_tables(const NodeArgs& a, uint32_t nc, const uint32_t* kept, const uint32_t* rank_kept,
                        uint32_t* oa, uint32_t* ob, uint32_t* ok, const uint32_t* rank_src,
                        const uint32_t* rank_snk, uint32_t n_src, uint32_t* onode, double* ogap, int32_t* oside,
                        hipStream_t s);
void launch_node_count(const NodeArgs& a, uint32_t nc, unsigned long long* out, hipStream_t s);

// ---- network extraction (nf_net.hip) ----
enum NetSummary {
  SUM_EDGES = 0,
  SUM_COMPONENTS = 1,
  SUM_LARGEST = 2,
  SUM_PERCOLATING = 3,
  SUM_NNZ = 4,
  SUM_SOURCE = 5,
  SUM_SINK = 6,
  NET_SUMMARY = 8
};

struct NetArgs {
  // contacts: CSR over the lower id, J > row, image applies to J
  const uint32_t* rowptr;
  const uint32_t* J;
  const int32_t* img;
  const double* gap64;
  const float4* P;
  const float4* D;
  uint32_t n;
  int axis;
  double L, gcut, lam, g0;
  // work arrays
  uint32_t* par;       // n: union-find parents, final component labels
  uint32_t* lowcnt;    // n + 1
  uint32_t* lowstart;  // n + 1
  uint32_t* lowfill;   // n
  uint32_t* lcol;      // lower-triangle entries (<= contacts)
  int32_t* limg;
  double* lg;
  uint32_t* fsrc;      // n + 1: electrode coupling flags
  uint32_t* fsnk;
  uint32_t* psrc;      // n + 1: their exclusive scans ([n] = count)
  uint32_t* psnk;
  double* gsrc;        // n: electrode conductances
  double* gsnk;
  uint32_t* csize;     // n: component sizes (at the root)
  uint32_t* cflag;     // n: component electrode flags (at the root)
  uint32_t* rowcnt;    // n + 2
  uint32_t* nrow;      // n + 3: CSR row pointers of the N + 2 node matrix
  unsigned long long* summary;  // NET_SUMMARY entries
  uint32_t* tmp;       // scan scratch
  // outputs of the fill phase
  uint32_t* ncol;
  double* nval;
  double* cg;    // n: compacted electrode conductances
  double* bsum;  // n / 1024 + 1: their block sums
};

// Components, couplings, lower lists, row counts, row pointers, summary.
void launch_network_build(const NetArgs& a, hipStream_t s);
// Row entries (ncol/nval sized from summary[SUM_NNZ]).
void launch_network_fill(const NetArgs& a, uint32_t n_src, uint32_t n_snk, hipStream_t s);

}  // namespace nf
