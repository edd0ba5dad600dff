/*
 * nanofiller.h - C ABI of the MI355X NanoFiller rod-packing library.
 *
 * Plain C types only (no HIP or torch types): callers bind it from C, C++,
 * Python ctypes or any FFI.  Every entry point returns an int status
 * (NF_OK = 0, negative NF_E_* on failure) unless stated otherwise; a message
 * for the last failure is available from nf_last_error().
 *
 * Units are micrometres.  `radius` is the rod (capsule) radius a: the
 * reference stores the same quantity in CNT_t::a (inc/CNT.h:10) and uses
 * a1+a2 as the hard-core distance (inc/intersection_dev.h:78).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * pyrlinum/NanoFiller).  The reference has no FFI of its own: its drivers are
 * in-process C++ calls sharing a simPrms object plus CUDA global state.
 */
#ifndef NANOFILLER_H
#define NANOFILLER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define NF_OK 0
#define NF_E_INVALID (-1)  /* bad argument or parameter value              */
#define NF_E_NOMEM (-2)    /* host or device allocation failed             */
#define NF_E_HIP (-3)      /* HIP runtime error                            */
#define NF_E_IO (-4)       /* file could not be read or written            */
#define NF_E_OVERFLOW (-5) /* a fixed-capacity cell or buffer overflowed   */
#define NF_E_STATE (-6)    /* call out of order (e.g. contacts before rods)*/
#define NF_E_NODEVICE (-7) /* no HIP device visible                        */
#define NF_E_CAPACITY (-8) /* caller buffer too small (size reported back) */

/* ---- laws --------------------------------------------------------------- */
enum { NF_LEN_CONST = 0, NF_LEN_LOGNORMAL = 1 };
enum { NF_ORT_ISO = 0, NF_ORT_VMF = 1 };
enum { NF_DEN_UNIFORM = 0, NF_DEN_AGGLOM = 1, NF_DEN_MESH = 2 };
enum { NF_PREC_FP64 = 0, NF_PREC_FP32_RECHECK = 1 };
enum { NF_STATUS_RUNNING = 0, NF_STATUS_TARGET = 1, NF_STATUS_MAX_ROUNDS = 2,
       NF_STATUS_STALLED = 3, NF_STATUS_OVERFLOW = 4 };

/* Generation parameters of one packing (one instance).
 * Mirrors the keyword-file state of simPrms (inc/simPrms.h:57-178) with the
 * additions the MI355X generator needs (seed, length/orientation/density
 * laws, schedule).  Fill with nf_params_default(), then nf_params_load() or
 * field assignment, then nf_params_normalize(). */
typedef struct nf_params {
  double box[3];        /* periodic box edges; normalised to float-exact   */
  uint32_t n_rods;      /* target rod count N (0: from vol_frac)           */
  uint32_t seed;        /* Philox4x32-10 key word 0                        */
  uint32_t stream;      /* Philox4x32-10 key word 1 (instance id)          */
  int32_t len_law;      /* NF_LEN_*                                        */
  double len;           /* constant length, or log-normal median          */
  double len_sigma;     /* log-normal sigma of ln L                        */
  double len_min;       /* log-normal clip range                           */
  double len_max;
  double radius;        /* rod radius a                                    */
  double sep;           /* INC_SEPARATION: extra hard-core distance        */
  double soft_frac;     /* INC_SOFTCORE: hard core = soft_frac*(a1+a2)+sep */
  int32_t ort_law;      /* NF_ORT_*                                        */
  double kappa;         /* von Mises-Fisher concentration                  */
  double axis[3];       /* vMF mean direction (normalised)                 */
  int32_t den_law;      /* NF_DEN_*                                        */
  int32_t n_agg;        /* number of Gaussian agglomerates                 */
  double agg_sigma;     /* agglomerate Gaussian sigma                      */
  double agg_peak;      /* peak density relative to background (>= 1)      */
  double vol_frac;      /* VOLUME_PERCENT / 100 (used when n_rods == 0)     */
  double contact_gap;   /* STAT_DIST_LIM: contact if surface gap <= this   */
  uint32_t max_rounds;  /* round cap (reference: 1000, cudaMain.cu:113)    */
  uint32_t stall_rounds;/* stop after this many rounds with no acceptance  */
  uint32_t min_batch;   /* candidates per round, lower clamp               */
  uint32_t max_batch;   /* candidates per round, upper clamp (0: derived, three excluded volumes' worth) */
  double cell_size;     /* generation cell edge; 0: auto (about one typical
                           pair reach, smaller for long rods in small boxes);
                           < 0: |cell_size| times the auto edge              */
  int32_t precision;    /* NF_PREC_*                                       */
  int32_t elec_axis;    /* network electrode axis 0/1/2 (x/y/z)            */
  double tunnel_lambda; /* tunnelling decay length (um), g = g0 exp(-gap/l)*/
  double tunnel_g0;     /* contact conductance prefactor                   */
  int32_t write_state;  /* WRT_BIN: the CLI saves output.bin after generate */
  int32_t den_ext[3];   /* NF_DEN_MESH: mesh points per axis                 */
  const float* den_mesh;/* NF_DEN_MESH: ext0*ext1*ext2 values in [0, 1], C
                           order (x slowest, as the reference's file); not
                           owned, read by nf_create                          */
} nf_params;

/* Generation statistics (reference prints these; cudaMain.cu:98-157). */
typedef struct nf_gen_stats {
  uint32_t n_target;      /* N requested                                   */
  uint32_t n_accepted;    /* rods in the final packing                     */
  uint32_t rounds;        /* insert/reject rounds run                      */
  int32_t status;         /* NF_STATUS_*                                   */
  uint64_t candidates;    /* Philox candidate slots drawn                  */
  uint64_t proposals;     /* candidates that passed the density test       */
  uint64_t pair_tests;    /* capsule predicate evaluations (pairs in reach)*/
  uint64_t rechecks;      /* FP32 results inside the error band, redone FP64*/
  uint64_t rejected_old;  /* proposals overlapping an accepted rod         */
  uint64_t rejected_new;  /* proposals overlapping a higher-priority one   */
  double wall_ms;         /* host clock, first candidate -> packing on HBM */
  double kern_ms[8];      /* HIP-event time per kernel class, see nf_kern  */
  double max_abs_d32_d64; /* max |d32-d64| over full tests (diagnostic run)*/
  uint64_t tile_bytes;    /* bytes the reject-vs-accepted kernel streamed
                             (probes and neighbour tiles; roofline figure) */
  uint64_t tile_bytes_new;/* same for the reject-vs-same-round kernel      */
  uint64_t rounds_launched; /* rounds launched (incl. no-op tail rounds)   */
  uint32_t slot_growths;  /* times the cell slots of the accepted rods were
                             doubled and the packing generated again (a
                             cell held more rods than planned)              */
  uint32_t reserved;
} nf_gen_stats;

/* Kernel classes reported in nf_gen_stats.kern_ms. */
enum { NF_K_GEN = 0, NF_K_BIN = 1, NF_K_REJECT_OLD = 2, NF_K_REJECT_NEW = 3,
       NF_K_COMMIT = 4, NF_K_CONTROL = 5, NF_K_CONTACT = 6, NF_K_OTHER = 7 };

typedef struct nf_handle nf_handle; /* opaque; one HIP stream per handle */

/* ---- parameters --------------------------------------------------------- */

/* Defaults of simPrms::simPrms (src/simPrms.cpp:22-117), with the target
 * held fixed (SELF_ALLIGN_STEP 0) and VOL_FRAC_PRECISION 0. */
void nf_params_default(nf_params* p);

/* Keyword parameter file, same keywords as parMap (inc/parameter_input.h:39-69)
 * and simPrms::set_ParamsFromFile (src/simPrms.cpp:212-327), plus the
 * keywords listed in the file's header comment. */
int nf_params_load(const char* path, nf_params* p);

/* Derived values: N from the volume fraction as simPrms::set_numCNT
 * (src/simPrms.cpp:987-1002), float-exact box edges, batch bounds. */
int nf_params_normalize(nf_params* p);

/* Centres of the Gaussian agglomerates of the density law (NF_DEN_AGGLOM):
 * 3 doubles each, as the generator places them (a pure function of seed,
 * stream and box).  Returns the count (0 for other laws), or a negative
 * status; out may be NULL to ask for the count.  The density is then
 * rho(x) = 1 + (agg_peak - 1) exp(-d^2 / (2 agg_sigma^2)), d the periodic
 * distance to the nearest centre, replacing the reference's density texture
 * (src/den2tex.h:5-53) for analytic profiles. */
int nf_agglomerate_centres(const nf_params* p, double* out, int32_t cap);

/* ---- lifecycle ---------------------------------------------------------- */

/* One packing.  Replaces the simPrms set-up in cudaMain (src/cudaMain.cu:61-97). */
int nf_create(const nf_params* p, nf_handle** out);
/* n independent packings generated together (batched persistent rounds). */
int nf_create_batch(const nf_params* ps, int n, nf_handle** out);
void nf_destroy(nf_handle* h);
/* The normalised parameters of instance inst (no counterpart in the
 * reference, whose simPrms fields are read directly; e.g. the box of a
 * handle made by nf_expand). */
int nf_get_params(nf_handle* h, int inst, nf_params* out);
/* Message of the last failure on h (h == NULL: last failure of nf_create*). */
const char* nf_last_error(const nf_handle* h);
/* Number of instances held by h. */
int nf_num_instances(const nf_handle* h);

/* ---- generation: cudaMain (src/cudaMain.cu:61-161) ---------------------- */

/* Runs the insert/reject rounds for every instance of h until each reaches
 * its target, its round cap or stalls.  st (may be NULL) receives one record
 * per instance. */
int nf_generate(nf_handle* h, nf_gen_stats* st);

/* Rods of instance inst in id order, 8 floats each: x y z cx cy cz l a
 * (the CNT_t record of inc/CNT.h:6-12 without the diagnostic k).
 * *n receives the rod count; NF_E_CAPACITY if cap < count. */
int nf_get_rods(nf_handle* h, int inst, float* rods, uint32_t cap, uint32_t* n);

/* Replaces the packing of instance inst with caller rods (state reload,
 * read_state_BIN in src/stateSR.cpp:95-134). */
int nf_set_rods(nf_handle* h, int inst, const float* rods, uint32_t n);

/* ---- contacts: get_contact_counts / get_contact_arrays
 *      (src/get_contact_arrays.cu:12-140) and get_contact_distance
 *      (src/get_contact_data.cu:71-128) ----------------------------------- */

/* Finds every pair (I < J, periodic image) whose surface gap is <= gap on
 * the post-processing instance (NF_OPT_POST_INSTANCE, default 0).  Result
 * stays on the device; *n_contacts receives its size. */
int nf_contacts(nf_handle* h, double gap, uint64_t* n_contacts);

/* Copies the contact list, sorted by (I, J, image): row_ptr has N+1 entries
 * (CSR over I), col/gap/image have n_contacts entries.  gap is the surface
 * gap d - a_I - a_J rounded from FP64; image packs the periodic image of J
 * as three signed bytes (x | y<<8 | z<<16).  Any pointer may be NULL. */
int nf_get_contacts(nf_handle* h, uint32_t* row_ptr, uint32_t* col, float* gap,
                    int32_t* image, uint64_t cap);

/* FP64 surface gaps of the same contacts, in the same order (the value the
 * FP32 gap is rounded from; the conductances of nf_network use it).  The
 * reference keeps only an FP32 distance (src/get_contact_data.cu:28-68). */
int nf_get_contact_gap64(nf_handle* h, double* gap64, uint64_t cap);

/* Contact (tunnelling) areas S of the same contacts, in the same order:
 * get_contact_surface / get_surface_cyl2cyl (src/get_contact_data.cu:
 * 130-217) with the area model of surf_to_cylinder_capped (inc/intersect/
 * intersect_cylinder_capped.h:204-293): the larger of the point-contact disc
 * pi dr^2, the crossing ellipse pi/2 dr (A+ + A-) and the near-parallel strip
 * 2 dr |overlap|, dr the radius of the surface patch within the contact gap
 * of the other rod.  FP64, computed on the device with the contact list. */
int nf_get_contact_area(nf_handle* h, double* area, uint64_t cap);

/* ---- histograms: statKernel_ANGL / statKernel_MUT / centre profile
 *      (src/statistics_collection.cu:640-861, inc/statKernel_ANGL.h:13-76,
 *      inc/statKernel_MUT.h:26-131) ---------------------------------------- */

typedef struct nf_hist_spec {
  int32_t bins;       /* bins per 1-D histogram                           */
  int32_t centre_axis;/* centre-density histogram along this axis          */
  double pol_axis[3]; /* polar angle measured from this direction          */
  int32_t phi_bins;   /* phi-theta 2-D histogram: phi bins (0: skip)       */
  int32_t theta_bins; /* phi-theta 2-D histogram: theta bins               */
} nf_hist_spec;

/* Exact int64 counts.  out_centre[bins], out_polar[bins], out_mutual[bins]
 * (mutual angle of the pairs of the last nf_contacts call), out_phitheta
 * [phi_bins*theta_bins] (may be NULL when phi_bins == 0). */
int nf_histograms(nf_handle* h, const nf_hist_spec* spec, int64_t* out_centre,
                  int64_t* out_polar, int64_t* out_mutual, int64_t* out_phitheta);

/* nf_histograms without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, limits, errors and results as nf_histograms; its scratch
 * (bin edges and counters, at most 64 KiB) lives in a static device array,
 * one per GPU, held until the process ends.  Calls that overlap on one GPU
 * (two threads, or two handles' streams) race on that array. */
int nf_histograms_noalloc(nf_handle* h, const nf_hist_spec* spec, int64_t* out_centre,
                          int64_t* out_polar, int64_t* out_mutual, int64_t* out_phitheta);

/* Centre-centre distance statistics, statKernel_CENT (inc/statKernel_CENT.h:
 * 8-122) driven by collectStats (src/statistics_collection.cu:866-1033), on
 * the post-processing instance: every unordered pair of distinct rods over all periodic images
 * whose centre distance d is below bins * width (at most 8 box edges).
 * counts[k] (int64, exact) holds d in [k width, (k+1) width); *n_pairs the
 * pair count; *sum_d_q24 and *sum_d2_q16 the exact sums of round(d * 2^24)
 * and round(d^2 * 2^16) (mean = sum_d_q24 / 2^24 / n_pairs; the reference's
 * float accumulators of d and d^2).  bins <= 1024.  Any pointer may be NULL. */
int nf_centre_distance(nf_handle* h, int32_t bins, double width, int64_t* counts,
                       uint64_t* n_pairs, uint64_t* sum_d_q24, uint64_t* sum_d2_q16);

/* Van der Waals energy, statKernel_VdW_naive (inc/statKernel_MUT.h:531-628)
 * driven by collect_vdw (src/statistics_collection.cu:53-154), over the
 * contacts of the last nf_contacts call with dist_limit = that call's gap:
 * pairs with interior closest points and gap < dist_limit contribute
 * excA / theta / a1 * curve2surface_limit(gap, dist_limit, a1, a2, A, B,
 * theta) (lengths in Angstrom; the reference's constants are A = 15.2,
 * B = 24100).  FP64, summed in a fixed order. */
int nf_vdw(nf_handle* h, double A, double B, double* energy, uint64_t* n_pairs);

/* ---- network: collect_cunduct_data (src/conductivity.cu:140-423) and
 *      adjacency_graph (src/adjacency_graph.cu:57-219) --------------------- */

typedef struct nf_network_info {
  uint32_t n_nodes;        /* N + 2 (rods, source electrode, sink)          */
  uint32_t n_components;   /* connected components of the rod graph          */
  uint32_t largest;        /* rods in the largest component                  */
  uint32_t percolating;    /* components touching both electrodes           */
  uint64_t nnz;            /* CSR non-zeros of the Kirchhoff matrix          */
  uint64_t n_edges;        /* rod-rod conductances (contacts not crossing the
                              electrode axis boundary)                       */
  uint32_t n_source;       /* rods coupled to the source electrode           */
  uint32_t n_sink;         /* rods coupled to the sink electrode             */
} nf_network_info;

/* Builds connected components and the FP64 CSR Kirchhoff (graph Laplacian)
 * matrix from the contacts of the last nf_contacts call.  The electrode axis
 * (params.elec_axis) is treated as non-periodic with electrodes at 0 and at
 * the box edge; rods within the contact gap of a plane couple to it. */
int nf_network(nf_handle* h, nf_network_info* info);

/* Copies the network: labels[N] (component = smallest rod id in it),
 * row_ptr[N+3], col[nnz], val[nnz] (FP64).  Any pointer may be NULL. */
int nf_get_network(nf_handle* h, uint32_t* labels, uint32_t* row_ptr,
                   uint32_t* col, double* val, uint64_t cap);

/* Dead-end pruning, the `reduce` verb (src/NanoFiller.cpp:71-90) with
 * reduce() (src/conductivity.cu:79-138) and contact_count_kernel_full /
 * reductKernel (inc/conductKernels.h:40-162): rounds of simultaneous removal
 * of every rod with fewer than min_cont contacts among the rods still
 * present (the reference uses 2), until a round removes nothing.  Contacts
 * are those of the last nf_contacts call, one per periodic image, without
 * those crossing the electrode axis (params.elec_axis); a coupling to an
 * electrode plane (the rule of nf_network) counts as one contact - the
 * reference instead keeps every rod of its electrode cell layers.  The kept
 * rods replace the post-processing instance's (NF_OPT_POST_INSTANCE, default 0) packing in id order (new ids 0..n_after-1); the
 * contact list and everything built on it are released, so nf_contacts
 * must run again for the pruned packing. */
typedef struct nf_reduce_info {
  uint32_t n_before;  /* rods before pruning                                 */
  uint32_t n_after;   /* rods kept                                           */
  uint32_t rounds;    /* removal rounds, the last one (removing nothing) included,
                         as the reference's iteration count                  */
  uint32_t pad;
  double device_ms;   /* device time of the rounds and the compaction        */
} nf_reduce_info;
int nf_reduce(nf_handle* h, int32_t min_cont, nf_reduce_info* info);
/* Old id of every rod kept by the last nf_reduce (n_after entries). */
int nf_get_reduce_kept(nf_handle* h, uint32_t* old_ids, uint64_t cap);

/* DC solve of the network between the electrodes (source at potential 1,
 * sink at 0): replaces tools/kirchhoff/kirchhoff.py:256-272 (direct sparse
 * solve of diag(S) - A - A^T with the source couplings as right-hand side,
 * current I = sum (1 - V_i) B_i) for the real conductances of nf_network. */
typedef struct nf_conductance_info {
  double conductance;   /* I = current from the source (units of tunnel_g0)   */
  double current_sink;  /* current into the sink; equals I at convergence     */
  double residual;      /* final ||b - A x|| / ||b|| of the reduced system    */
  double solve_ms;      /* device time of set-up-free CG iterations + currents */
  uint32_t unknowns;    /* rods in components touching both electrodes         */
  uint32_t iterations;  /* conjugate-gradient iterations                       */
  int32_t converged;    /* 1 when residual <= tol                              */
  uint32_t pad;
} nf_conductance_info;

/* Jacobi-preconditioned FP64 conjugate gradients on the rods of the
 * components that touch both electrodes (the only ones carrying current),
 * until ||r|| <= tol ||b|| or max_iter iterations.  potentials (N doubles,
 * may be NULL) receives every rod's potential: solved, 1 or 0 for components
 * touching only the source or only the sink, NaN for floating ones.
 * Requires nf_network. */
int nf_conductance(nf_handle* h, double tol, uint32_t max_iter, nf_conductance_info* info,
                   double* potentials);

/* Node-level segment network: the internal-resistance branch of
 * collect_cunduct_data (src/conductivity.cu:213-410): contact points as
 * nodes, segments along the rods, screened contacts removed
 * (check_screened_contacts :430-528), nodes within one radius merged
 * (remove_redundant_nodes / reduce_node_tbl, src/adjacency_graph.cu:57-219).
 * Built from the contacts of the last nf_contacts call and the electrode
 * rules of nf_network (params.elec_axis, contact_gap). */
typedef struct nf_node_info {
  uint32_t n_points;    /* contact and electrode points before merging      */
  uint32_t n_nodes;     /* nodes after merging                              */
  uint32_t n_segments;  /* segments between consecutive nodes of a rod      */
  uint32_t n_edges;     /* rod-rod contacts kept (network, not screened)   */
  uint32_t n_screened;  /* contacts removed: a third rod blocks them       */
  uint32_t n_source;    /* electrode points on the source plane            */
  uint32_t n_sink;      /* electrode points on the sink plane              */
  uint32_t pad;
  double device_ms;     /* device time of the build                        */
} nf_node_info;

int nf_node_network(nf_handle* h, nf_node_info* info);
/* Nodes in rod order, then t: rod[n_nodes], t[n_nodes] (axial position of
 * the node, length units, mean of its merged points). */
int nf_get_nodes(nf_handle* h, uint32_t* rod, double* t, uint64_t cap);
/* Segments: node a < node b on one rod, L = (t_b - t_a) / radius, A = radius
 * (get_segment_length, src/get_contact_data.cu:235-297). */
int nf_get_segments(nf_handle* h, uint32_t* na, uint32_t* nb, double* L, double* A, uint64_t cap);
/* Kept rod-rod contacts as node pairs, in contact-list order; contact[]
 * gives each one's index in the nf_get_contacts arrays.  All pointers NULL:
 * count only. */
int nf_get_node_edges(nf_handle* h, uint32_t* na, uint32_t* nb, uint32_t* contact, uint64_t cap,
                      uint64_t* count);
/* Electrode couplings as (node, surface gap, side 0 = source / 1 = sink):
 * all source couplings in rod order, then the sink ones. */
int nf_get_node_electrodes(nf_handle* h, uint32_t* node, double* gap, int32_t* side, uint64_t cap,
                           uint64_t* count);

/* ---- output writers ---------------------------------------------------- */

/* Legacy-ASCII VTK of the post-processing instance as write_CNTvtk (src/IO.cpp:38-76). */
int nf_write_vtk(nf_handle* h, const char* path);
/* Contact lists as save_coo (src/conductivity.cu:838-879): <prefix>internal.srt.coo,
 * <prefix>sink.elec.coo, <prefix>source.elec.coo. */
int nf_write_coo(nf_handle* h, const char* prefix);
/* Node network files of the internal-resistance branch (save_coo,
 * src/conductivity.cu:881-930): <prefix>Nodes.tbl.dat (node, rod, t),
 * segment.LA.coo (a, b, L, A), internal.DS.coo (a, b, D, S), sink.DS.coo
 * and source.DS.coo (electrode id, node, D, S). */
int nf_write_node_coo(nf_handle* h, const char* prefix);

/* ---- density mesh files ------------------------------------------------ */

/* Reads a density map in the reference's text format (read_density,
 * src/IO.cpp:11-36): physDim[3], ext[3], then ext0*ext1*ext2 values in C
 * order ([ix][iy][iz], x slowest).  values may be NULL (header only); cap
 * is its length.  The density at a point is the trilinear interpolation of
 * the map with periodic wrap and mesh points at the centres of ext cells of
 * the box, as the reference's normalised, linearly filtered, wrapping 3-D
 * texture (src/den2tex.h:5-53) - evaluated exactly in FP64. */
int nf_read_density_mesh(const char* path, double dim[3], int32_t ext[3], float* values,
                         uint64_t cap);
/* Writes a density map in the same format. */
int nf_write_density_mesh(const char* path, const double dim[3], const int32_t ext[3],
                          const float* values);

/* ---- state files ------------------------------------------------------- */

/* Header of the reference's binary state file (output.bin, state_<vf>.bin;
 * save_state_BIN / read_state_BIN, src/stateSR.cpp:54-134). */
typedef struct nf_state_header {
  float box[3];     /* physDim                                      */
  float edge[3];    /* cell edges                                   */
  int32_t ext[3];   /* cells per axis                               */
  float len;        /* default length l                             */
  float radius;     /* default radius a                             */
  uint32_t slots;   /* ttlCNT: slots in the file                    */
  uint32_t n_rods;  /* numCNT                                       */
  float tol;        /* TOL: slots with l <= tol are empty           */
} nf_state_header;

/* Writes the post-processing instance as save_state_BIN after simPrms::repack
 * (src/stateSR.cpp:54-88, src/simPrms.cpp:691-743): the generation cell
 * grid, per-cell occupancy and slot prefix, and the nine field-major
 * arrays x y z cx cy cz l a k (k = cell index), one slot per rod. */
int nf_save_state(nf_handle* h, const char* path);
/* Reads the header of a state file and checks the file size against it. */
int nf_read_state_header(const char* path, nf_state_header* out);
/* Replaces the post-processing instance's (NF_OPT_POST_INSTANCE, default 0) rods with the state file's (read_state_BIN,
 * src/stateSR.cpp:90-134): every slot with l > TOL, in slot order, as
 * simPrms::asmbl_cnt_data (src/simPrms.cpp:329-391).  The file's box must
 * equal the handle's; the rod count must not exceed the handle's target. */
int nf_load_state(nf_handle* h, const char* path, uint32_t* n_rods);

/* Milestone state dumps during nf_generate (VOLUME_PERCENT_LIST,
 * src/cudaMain.cu:136-142, listed at src/simPrms.cpp:246-251): after each
 * round whose packing has reached the next listed volume fraction (percent,
 * N pi a^2 l / V * 100 with the nominal a and l), the packing is written as
 * nf_save_state to <dir>/state_<vf>.bin, vf the fraction reached ("%4.2f"),
 * and every listed value reached by then is passed, as the reference's one
 * check per round does.  The packing is the same as without milestones; the
 * rounds are synchronised with the host one at a time while a milestone is
 * pending.  n = 0 clears the list.  Single-instance handles. */
int nf_set_milestones(nf_handle* h, const double* vf_percent, int32_t n, const char* dir);
/* The VOLUME_PERCENT_LIST of a parameter file ("VOLUME_PERCENT_LIST n v1 ..
 * vn", percent): returns n (0 when the keyword is absent) and fills up to
 * cap values. */
int nf_params_milestones(const char* path, double* vf_percent, int32_t cap);

/* expand (simPrms::expand, src/simPrms.cpp:744-852; verb at
 * src/NanoFiller.cpp:67-69): a new single-instance handle *out holding
 * the post-processing instance's (NF_OPT_POST_INSTANCE, default 0) packing with its box grown along axis k = dir >> 1 (dir the
 * reference's bit mask: 1 = x, 2 = y, 4 = z) by a buffer of periodic copies
 * on each side.  The buffer is B = ceil((l_max + contact_gap) / E) E, E the
 * generation cell edge along k (the reference's NeiOrder cells of the
 * domain grid); it must not exceed the box.  Copies of the rods whose centre
 * lies within B of the far face go below the box, copies of those within B
 * of the near face above it, and every coordinate along k is shifted by B so
 * that the new box is [0, L + 2B).  Rod i keeps id i; the copies follow in id
 * order, first those below, then those above.  h is left unchanged. */
int nf_expand(nf_handle* h, int dir, nf_handle** out);

/* ---- library ----------------------------------------------------------- */

/* Version string of the library build. */
const char* nf_version(void);

/* Handle options (no counterpart in the reference, which has no runtime
 * switches beyond compile-time flags). */
enum {
  NF_OPT_GRAPH = 1,        /* 1: replay generation rounds from a HIP graph (default 0) */
  NF_OPT_PROFILE = 2,      /* 1: time each kernel class with HIP events (no graph)   */
  NF_OPT_DIAG = 3,         /* 1: also compute FP64 for FP32-decided pairs and report
                              max |d32 - d64|                                         */
  NF_OPT_CONTACT_PREC = 4, /* NF_PREC_* for nf_contacts (-1: the params' precision)   */
  NF_OPT_CONTACT_ENGINE = 5, /* NF_ENGINE_* candidate pairs of nf_contacts (default
                                footprint)                                                */
  NF_OPT_FOOTPRINT_FINE = 6, /* footprint engine: fine cell edge in 1/1000 of the typical
                                rod length (0: default 200)                               */
  NF_OPT_FOOTPRINT_OCC = 7,  /* footprint engine: rods crossing a home cell at the mean
                                density, which sizes the home cells (0: default 64)       */
  NF_OPT_POST_INSTANCE = 8,  /* the instance of a batch that contacts, statistics,
                                network, reduce, conductance, the writers, the state
                                files and expand act on (default 0); changing it drops
                                the results computed for the previous one              */
  NF_OPT_SLOT_SCALE = 9,     /* per-cell slots of the accepted rods in 1/1000 of the
                                planned capacity (default 1000); a generation whose
                                slots overflow is repeated with twice as many, with
                                the same packing (nf_gen_stats.slot_growths)           */
  /* Launch shape of the generation's pair passes (the packing does not
   * depend on it; nanofiller_amd/autotune.py picks it per workload): */
  NF_OPT_WAVES_OLD = 10,     /* waves per CU, reject vs accepted (0: default 16)     */
  NF_OPT_WAVES_NEW = 11,     /* waves per CU, reject vs same-round (0: default 16)   */
  NF_OPT_PART_SHIFT = 12,    /* reject vs accepted: a dense cell is dealt over
                                1 << shift work items in launches with few cells per
                                wave (-1: default 1)                                    */
  NF_OPT_SPLIT_MIN = 13,     /* probes per work item of a dealt cell, at least
                                (0: default 32)                                         */
  NF_OPT_TEST_ALL_PAIRS = 14 /* 1: the reject-vs-accepted pass tests every pair within
                                reach, so that the pair-test count is the oracle's;
                                0 (default): a proposal leaves its stencil walk once an
                                accepted rod cuts it (same decisions, fewer tests)      */
};
/* Candidate pairs of nf_contacts.  Both give the same contact list.
 * FOOTPRINT: each rod binned into the fine cells its capsule (radius
 * a + gap/2) can touch; rods sharing a cell are tested, each contact
 * reported by the cell holding the midpoint of its closest points.
 * CELLS: the cell-pair engine of the generator over a centre cell list; its
 * capsule-test count is the number of pairs within centre reach (the
 * reference's stencil count). */
enum { NF_ENGINE_FOOTPRINT = 0, NF_ENGINE_CELLS = 1 };
int nf_set_option(nf_handle* h, int key, int64_t value);

/* Statistics of the last nf_contacts call: capsule tests on pairs in reach
 * (with the footprint engine: on the candidate pairs within centre reach, a
 * pair counted once per cell the two rods share),
 * FP64 rechecks, max |d32-d64| (NF_OPT_DIAG), HIP-event times in ms of the
 * four contact stages (binning, the pair pass that finds and appends every
 * contact, CSR build, row sort; kernel_ms[4]) and the probe + tile bytes the
 * pair pass streamed. */
int nf_contacts_stats(nf_handle* h, uint64_t* pair_tests, uint64_t* rechecks, double* max_dd,
                      float* kernel_ms, uint64_t* tile_bytes);

/* Engine that found the last contact list (NF_ENGINE_*; the footprint
 * engine falls back to the cell engine when a single home cell is crossed
 * by more rods than fit), and for the footprint engine its home cell edge
 * (um), fine cells per axis (0 for the cell engine) and the number of work
 * units (runs of home cells) it redid one cell at a time.  No counterpart in
 * the reference: a diagnostic for tests and tuning. */
int nf_contacts_engine(nf_handle* h, int* engine, double* home_edge, int* fine, uint64_t* split_units);

/* Diagnostics (no counterpart in the reference): evaluates one FP64 device
 * building block element-wise on the GPU so callers can check it against
 * the CPU.  which: 0 correctly rounded sqrt, 1 exp, 2 log, 3 division
 * (in holds 2n values, out[i] = in[2i]/in[2i+1]), 4 Philox-word-to-uniform
 * conversion (bits of in[i] = hi:lo words), 5 the device's own sqrt, 6 a
 * plain copy (out[i] = in[i]; a known-size kernel for counter calibration). */
int nf_selftest(int which, const double* in, double* out, uint64_t n);

/* Diagnostics (no counterpart in the reference): the FP32 screen and band of
 * the pair passes, evaluated on the GPU for n pairs, so that tests can check
 * the device's decisions (its hardware square root and reciprocal) against
 * FP64.  ra, rb: 8 floats per rod (centre, direction, length, radius);
 * shift: 3 doubles per pair, added to rod b's image; contact: 1 for the
 * contact cut (a_1 + a_2 + sep_or_gap), 0 for the generation's
 * (soft_frac (a_1 + a_2) + sep_or_gap); ew: the FP32 bound on a
 * coordinate-difference component.  Out per pair: screen (-1 undecided,
 * 0 out of reach, 1 far, 2 full test), decided (-1 undecided, else 1 cut,
 * 0 not cut) and within (reach, when decided). */
int nf_selftest_pairs(const float* ra, const float* rb, const double* shift, uint64_t n, int contact,
                      double sep_or_gap, double soft_frac, float ew, int32_t* screen, int32_t* decided,
                      int32_t* within);

#ifdef __cplusplus
}
#endif
#endif /* NANOFILLER_H */
