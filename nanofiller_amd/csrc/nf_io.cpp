// nf_io.cpp - keyword parameter files and the VTK / COO writers.
//
// Keywords of parMap (inc/parameter_input.h:39-69) with the semantics of
// simPrms::set_ParamsFromFile (src/simPrms.cpp:212-327), plus:
//   NUM_RODS n                       target rod count (overrides VOLUME_PERCENT)
//   RNG_SEED seed [stream]           Philox key
//   INC_LEN_LAW CONST | LOGNORMAL median sigma min max
//   ORT_LAW ISO | VMF kappa ax ay az
//   DENSITY_LAW UNIFORM | AGGLOM n sigma peak
//   SPAT_DISTR_FILE file             density map (read_density format); sets
//                                    the box and the mesh density law
//   CELL_SIZE edge                   generation cell edge (um)
//   PRECISION FP64 | FP32            FP32 = FP32 with FP64 re-check band
//   ELECTRODE_AXIS 0|1|2|x|y|z, TUNNEL_LAMBDA l, TUNNEL_G0 g0
//   MAX_ROUNDS n, STALL_ROUNDS n, MIN_BATCH n, MAX_BATCH n
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "nf_host.h"

using namespace nf;

static std::string upper(std::string s) {
  for (auto& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

extern "C" int nf_params_load(const char* path, nf_params* p) {
  if (!path || !p) return fail(nullptr, NF_E_INVALID, "null argument");
  std::ifstream f(path);
  if (!f.is_open()) return fail(nullptr, NF_E_IO, std::string("File ") + path + " not found");
  const double PI = 3.14159265358979323846;
  double theta_med = 0.0, phi_med = 0.0, theta_dev = -1.0, phi_dev = -1.0;
  double pref_phi = 0.0, pref_theta = 0.0;
  bool ort_explicit = false, ort_dev_given = false;
  std::string warn;
  std::string line;
  int lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    const size_t b = line.find_first_not_of(" \t\r");
    if (b == std::string::npos) continue;
    if (line[b] == '#') continue;
    std::istringstream is(line.substr(b));
    std::string key;
    is >> key;
    auto bad = [&](const char* what) {
      return fail(nullptr, NF_E_INVALID,
                  std::string(path) + ":" + std::to_string(lineno) + ": " + key + ": " + what);
    };
    if (key == "WRT_BIN") {
      int v = 0;
      if (!(is >> v)) return bad("expected 0 or 1");
      p->write_state = v != 0;
    } else if (key == "VOLUME_PERCENT") {
      double v;
      if (!(is >> v)) return bad("expected a number");
      p->vol_frac = v / 100.0;
      p->n_rods = 0;
    } else if (key == "VOLUME_PERCENT_LIST") {
      // milestone dumps: read by nf_params_milestones
    } else if (key == "VOL_FRAC_PRECISION") {
      double v;
      is >> v;
      if (v != 0.0) warn += "VOL_FRAC_PRECISION ignored: the generator targets N exactly; ";
    } else if (key == "SPAT_DISTR_FILE") {
      std::string fn;
      is >> fn;
      if (fn == "NONE") {
        int ext[3];
        if (!(is >> p->box[0] >> p->box[1] >> p->box[2])) return bad("expected NONE Lx Ly Lz ex ey ez");
        is >> ext[0] >> ext[1] >> ext[2];
      } else {
        // density map file: box from its dimensions, values kept by the
        // library for the life of the process (the params point at them)
        double dim[3];
        int32_t ext[3];
        if (nf_read_density_mesh(fn.c_str(), dim, ext, nullptr, 0) != NF_OK)
          return bad(("cannot read density map: " + g_last_error).c_str());
        std::vector<float> v((size_t)ext[0] * ext[1] * ext[2]);
        if (nf_read_density_mesh(fn.c_str(), dim, ext, v.data(), v.size()) != NF_OK)
          return bad(("cannot read density map: " + g_last_error).c_str());
        static std::deque<std::vector<float>> kept;
        kept.push_back(std::move(v));
        for (int d = 0; d < 3; ++d) {
          p->box[d] = dim[d];
          p->den_ext[d] = ext[d];
        }
        p->den_mesh = kept.back().data();
        p->den_law = NF_DEN_MESH;
      }
    } else if (key == "PREF_ORT_PHI") {
      is >> pref_phi;
      ort_dev_given = true;
    } else if (key == "PREF_ORT_THETA") {
      is >> pref_theta;
      ort_dev_given = true;
    } else if (key == "DEF_THETA_MED") {
      is >> theta_med;
    } else if (key == "DEF_PHI_MED") {
      is >> phi_med;
    } else if (key == "DEF_THETA_DEV") {
      is >> theta_dev;
    } else if (key == "DEF_PHI_DEV") {
      is >> phi_dev;
    } else if (key == "INC_LEN_MED") {
      if (!(is >> p->len)) return bad("expected a number");
    } else if (key == "INC_RAD_MED") {
      if (!(is >> p->radius)) return bad("expected a number");
    } else if (key == "INC_SEPARATION") {
      is >> p->sep;
    } else if (key == "INC_SOFTCORE") {
      is >> p->soft_frac;
    } else if (key == "STAT_DIST_LIM") {
      is >> p->contact_gap;
    } else if (key == "CLUST_FLAG") {
      int v = 0;
      is >> v;
      if (v) warn += "CLUST_FLAG ignored (clustering is not part of this generator); ";
    } else if (key == "SELF_ALLIGN_STEP") {
      int v = 0;
      is >> v;
      if (v) warn += "SELF_ALLIGN_STEP ignored: targets are held fixed; ";
    } else if (key == "REDUCE_FLAG" || key == "CLUST_DIST_LIM" || key == "CLUST_ANGL_LIM" ||
               key == "SPAT_DIMENSIONS" || key == "ORNT_DISTR_FILE" || key == "IDIM_DISTR_FILE" ||
               key == "INC_LEN_DEV" || key == "INC_RAD_DEV" || key == "MATH_TOLERANCE" ||
               key == "DMN_CELL_MARGE") {
      // parsed and ignored, as in simPrms::set_ParamsFromFile
    } else if (key == "NUM_RODS") {
      unsigned long v;
      if (!(is >> v) || v == 0 || v > 0xFFFFFFFFul) return bad("expected a positive count");
      p->n_rods = (uint32_t)v;
    } else if (key == "RNG_SEED") {
      unsigned long s;
      if (!(is >> s)) return bad("expected a seed");
      p->seed = (uint32_t)s;
      unsigned long t;
      if (is >> t) p->stream = (uint32_t)t;
    } else if (key == "INC_LEN_LAW") {
      std::string law;
      is >> law;
      law = upper(law);
      if (law == "CONST") {
        p->len_law = NF_LEN_CONST;
      } else if (law == "LOGNORMAL") {
        p->len_law = NF_LEN_LOGNORMAL;
        if (!(is >> p->len >> p->len_sigma >> p->len_min >> p->len_max))
          return bad("expected median sigma min max");
      } else {
        return bad("unknown length law");
      }
    } else if (key == "ORT_LAW") {
      std::string law;
      is >> law;
      law = upper(law);
      ort_explicit = true;
      if (law == "ISO") {
        p->ort_law = NF_ORT_ISO;
      } else if (law == "VMF") {
        p->ort_law = NF_ORT_VMF;
        if (!(is >> p->kappa >> p->axis[0] >> p->axis[1] >> p->axis[2]))
          return bad("expected kappa ax ay az");
      } else {
        return bad("unknown orientation law");
      }
    } else if (key == "DENSITY_LAW") {
      std::string law;
      is >> law;
      law = upper(law);
      if (law == "UNIFORM") {
        p->den_law = NF_DEN_UNIFORM;
      } else if (law == "AGGLOM") {
        p->den_law = NF_DEN_AGGLOM;
        if (!(is >> p->n_agg >> p->agg_sigma >> p->agg_peak)) return bad("expected n sigma peak");
      } else {
        return bad("unknown density law");
      }
    } else if (key == "CELL_SIZE") {
      is >> p->cell_size;
    } else if (key == "PRECISION") {
      std::string v;
      is >> v;
      v = upper(v);
      if (v == "FP64")
        p->precision = NF_PREC_FP64;
      else if (v == "FP32")
        p->precision = NF_PREC_FP32_RECHECK;
      else
        return bad("expected FP64 or FP32");
    } else if (key == "ELECTRODE_AXIS") {
      std::string v;
      is >> v;
      v = upper(v);
      if (v == "0" || v == "X") p->elec_axis = 0;
      else if (v == "1" || v == "Y") p->elec_axis = 1;
      else if (v == "2" || v == "Z") p->elec_axis = 2;
      else return bad("expected x, y or z");
    } else if (key == "TUNNEL_LAMBDA") {
      is >> p->tunnel_lambda;
    } else if (key == "TUNNEL_G0") {
      is >> p->tunnel_g0;
    } else if (key == "MAX_ROUNDS") {
      is >> p->max_rounds;
    } else if (key == "STALL_ROUNDS") {
      is >> p->stall_rounds;
    } else if (key == "MIN_BATCH") {
      is >> p->min_batch;
    } else if (key == "MAX_BATCH") {
      is >> p->max_batch;
    } else {
      warn += "unrecognised keyword " + key + "; ";
    }
  }
  // Reference orientation keywords: a positive deviation selects a
  // preferred direction (genCosine, inc/cuda_CNTgenerate.h:262-270); it maps
  // to a vMF law about (theta_med, phi_med) with kappa = 1/dev^2.
  if (!ort_explicit && (theta_dev > 0.0 || phi_dev > 0.0)) {
    const double dev = (theta_dev > 0.0 ? theta_dev : phi_dev) / 180.0 * PI;
    const double th = theta_med / 180.0 * PI, ph = phi_med / 180.0 * PI;
    p->ort_law = NF_ORT_VMF;
    p->kappa = 1.0 / (dev * dev);
    p->axis[0] = std::sin(th) * std::cos(ph);
    p->axis[1] = std::sin(th) * std::sin(ph);
    p->axis[2] = std::cos(th);
  }
  (void)pref_phi;
  (void)pref_theta;
  (void)ort_dev_given;
  g_last_error = warn;
  return NF_OK;
}

// ===========================================================================
// Density maps: read_density (src/IO.cpp:11-36) - text: physDim[3], ext[3],
// then the values in C order ([ix][iy][iz], x slowest).
// ===========================================================================
extern "C" int nf_params_milestones(const char* path, double* vf_percent, int32_t cap) {
  if (!path) return fail(nullptr, NF_E_INVALID, "null argument");
  std::ifstream f(path);
  if (!f.is_open()) return fail(nullptr, NF_E_IO, std::string("File ") + path + " not found");
  std::string line;
  int32_t count = 0;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string key;
    if (!(is >> key) || upper(key) != "VOLUME_PERCENT_LIST") continue;
    int32_t n = 0;
    if (!(is >> n) || n < 0) return fail(nullptr, NF_E_INVALID, std::string(path) + ": VOLUME_PERCENT_LIST: expected a count");
    for (int32_t i = 0; i < n; ++i) {
      double v;
      if (!(is >> v)) return fail(nullptr, NF_E_INVALID, std::string(path) + ": VOLUME_PERCENT_LIST: expected " + std::to_string(n) + " values");
      if (vf_percent && i < cap) vf_percent[i] = v;
    }
    count = n;
  }
  return count;
}

extern "C" int nf_read_density_mesh(const char* path, double dim[3], int32_t ext[3], float* values,
                                    uint64_t cap) {
  if (!path || !dim || !ext) return fail(nullptr, NF_E_INVALID, "null argument");
  std::ifstream f(path);
  if (!f.is_open()) return fail(nullptr, NF_E_IO, std::string("cannot open density map ") + path);
  if (!(f >> dim[0] >> dim[1] >> dim[2] >> ext[0] >> ext[1] >> ext[2]))
    return fail(nullptr, NF_E_INVALID, std::string(path) + ": expected Lx Ly Lz ex ey ez");
  for (int d = 0; d < 3; ++d)
    if (!(dim[d] > 0.0) || ext[d] < 1 || ext[d] > 4096)
      return fail(nullptr, NF_E_INVALID, std::string(path) + ": bad dimensions or extents");
  const uint64_t n = (uint64_t)ext[0] * ext[1] * ext[2];
  if (!values) return NF_OK;
  if (cap < n) return fail(nullptr, NF_E_CAPACITY, std::string(path) + ": value buffer too small");
  for (uint64_t i = 0; i < n; ++i) {
    double v;
    if (!(f >> v))
      return fail(nullptr, NF_E_INVALID, std::string(path) + ": expected " + std::to_string(n) +
                                             " values, found " + std::to_string(i));
    if (!(v >= 0.0 && v <= 1.0))
      return fail(nullptr, NF_E_INVALID, std::string(path) + ": density values must lie in [0, 1] "
                                         "(von Neumann acceptance)");
    values[i] = (float)v;
  }
  return NF_OK;
}

extern "C" int nf_write_density_mesh(const char* path, const double dim[3], const int32_t ext[3],
                                     const float* values) {
  if (!path || !dim || !ext || !values) return fail(nullptr, NF_E_INVALID, "null argument");
  std::ofstream f(path, std::ios::out | std::ios::trunc);
  if (!f.is_open()) return fail(nullptr, NF_E_IO, std::string("cannot write ") + path);
  f << std::setprecision(9) << dim[0] << " " << dim[1] << " " << dim[2] << "\n";
  f << ext[0] << " " << ext[1] << " " << ext[2] << "\n";
  const uint64_t n = (uint64_t)ext[0] * ext[1] * ext[2];
  for (uint64_t i = 0; i < n; ++i) f << values[i] << ((i + 1) % (uint64_t)ext[2] ? " " : "\n");
  if (!f.good()) return fail(nullptr, NF_E_IO, std::string("write failed: ") + path);
  return NF_OK;
}

// ===========================================================================
// VTK: write_CNTvtk (src/IO.cpp:38-76) - legacy ASCII unstructured grid with
// centres as POINTS, directions as VECTORS CNT and (l, a, id) as SCALARS LAK.
// ===========================================================================
extern "C" int nf_write_vtk(nf_handle* h, const char* path) {
  if (!h || !path) return fail(h, NF_E_INVALID, "null argument");
  const uint32_t n = h->n_rods[h->post_inst];
  std::vector<float> rods((size_t)n * 8);
  uint32_t got = 0;
  int rc = nf_get_rods(h, h->post_inst, rods.data(), n, &got);
  if (rc != NF_OK) return rc;
  std::ofstream f(path, std::ios::out | std::ios::trunc);
  if (!f.is_open()) return fail(h, NF_E_IO, std::string("cannot write ") + path);
  f << "# vtk DataFile Version 2.0" << std::endl;
  f << "CUDA generated CNTs" << std::endl;
  f << "ASCII" << std::endl;
  f << "DATASET UNSTRUCTURED_GRID" << std::endl;
  f << "POINTS " << n << " float" << std::endl;
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = &rods[8 * (size_t)i];
    f << std::setw(12) << std::setprecision(6) << r[0] << std::setw(12) << std::setprecision(6) << r[1]
      << std::setw(12) << std::setprecision(6) << r[2] << std::endl;
  }
  f << "POINT_DATA " << n << std::endl;
  f << "VECTORS CNT float" << std::endl;
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = &rods[8 * (size_t)i];
    f << std::setw(15) << std::setprecision(6) << r[3] << std::setw(15) << std::setprecision(6) << r[4]
      << std::setw(15) << std::setprecision(6) << r[5] << std::endl;
  }
  f << "SCALARS LAK float 3" << std::endl;
  f << "LOOKUP_TABLE default" << std::endl;
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = &rods[8 * (size_t)i];
    f << std::setw(12) << std::setprecision(6) << r[6] << std::setw(12) << std::setprecision(6) << r[7]
      << std::setw(12) << std::setprecision(6) << (float)i << std::endl;
  }
  if (!f.good()) return fail(h, NF_E_IO, std::string("write failed: ") + path);
  return NF_OK;
}

// ===========================================================================
// COO: save_coo (src/conductivity.cu:838-879) - "#\tI\tJ\tD\tS" header,
// one "\tI\tJ\tD\tS" line per contact, default stream formatting.
// internal.srt.coo holds the rod-rod contacts of the network (image not
// crossing the electrode axis); source.elec.coo / sink.elec.coo hold the
// rod-electrode couplings with J = 4294967295 (source) / 4294967294 (sink),
// the SRC_ELEC / SNK_ELEC ids of inc/definitions.h:17-19.  D is the surface
// gap.  S of a rod-rod contact is its contact area from the device
// (nf_get_contact_area: point disc, crossing ellipse or parallel strip of
// surf_to_cylinder_capped, inc/intersect/intersect_cylinder_capped.h:
// 204-293); a rod-plane coupling has no second axis, so its S is the
// point-contact disc pi*dr^2 of that model's default branch (:224-233) with
// the rod's own radius on both sides.
// ===========================================================================
static double point_area(double gap, double a1, double a2, double rlim) {
  const double PI = 3.14159265358979323846;
  const double d = gap + a1 + a2;
  const double h1 = std::max(d - a2 - rlim, 0.0), h2 = std::max(d - a1 - rlim, 0.0);
  double dr1 = 0.0, dr2 = 0.0;
  if (h1 <= a1) dr1 = a1 * std::sqrt(1.0 - (h1 / a1) * (h1 / a1));
  if (h2 <= a2) dr2 = a2 * std::sqrt(1.0 - (h2 / a2) * (h2 / a2));
  const double dr = std::min(dr1, dr2);
  return PI * dr * dr;
}

extern "C" int nf_write_coo(nf_handle* h, const char* prefix) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  if (!h->cb.valid) return fail(h, NF_E_STATE, "nf_contacts must run before nf_write_coo");
  const std::string pre = prefix ? prefix : "";
  const ContactBuffers& c = h->cb;
  const uint32_t n = c.n_rods;
  const uint64_t C = c.n_contacts;
  std::vector<uint32_t> row(n + 1), col(C);
  std::vector<float> gap(C);
  std::vector<int32_t> img(C);
  int rc = nf_get_contacts(h, row.data(), col.data(), gap.data(), img.data(), C);
  if (rc != NF_OK) return rc;
  std::vector<float> rods((size_t)n * 8);
  uint32_t got = 0;
  rc = nf_get_rods(h, h->post_inst, rods.data(), n, &got);
  if (rc != NF_OK) return rc;
  std::vector<double> area(C);
  rc = nf_get_contact_area(h, area.data(), C);
  if (rc != NF_OK) return rc;
  const int ax = h->params[h->post_inst].elec_axis;
  const double rlim = c.gap_cut;
  {
    std::ofstream f(pre + "internal.srt.coo");
    if (!f.is_open()) return fail(h, NF_E_IO, "cannot write " + pre + "internal.srt.coo");
    f << "#\tI\tJ\tD\tS\n";
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t k = row[i]; k < row[i + 1]; ++k) {
        const int im = (int)(int8_t)((img[k] >> (8 * ax)) & 0xFF);
        if (im != 0) continue;
        f << "\t" << i << "\t" << col[k] << "\t" << gap[k] << "\t" << (float)area[k] << "\n";
      }
  }
  // electrode couplings: planes at 0 and box[ax]
  const double L = h->params[h->post_inst].box[ax];
  std::ofstream fk(pre + "sink.elec.coo"), fs(pre + "source.elec.coo");
  if (!fk.is_open() || !fs.is_open()) return fail(h, NF_E_IO, "cannot write electrode COO files");
  fk << "#\tI\tJ\tD\tS\n";
  fs << "#\tI\tJ\tD\tS\n";
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = &rods[8 * (size_t)i];
    const double ext = std::fabs(((double)r[6] * 0.5) * (double)r[3 + ax]);
    const double pos = (double)r[ax];
    const double gsrc = (pos - ext) - (double)r[7];
    const double gsnk = (L - (pos + ext)) - (double)r[7];
    if (gsrc <= rlim) {
      const double g = gsrc > 0.0 ? gsrc : 0.0;
      fs << "\t" << i << "\t" << 4294967295u << "\t" << (float)g << "\t"
         << (float)point_area(g, r[7], r[7], rlim) << "\n";
    }
    if (gsnk <= rlim) {
      const double g = gsnk > 0.0 ? gsnk : 0.0;
      fk << "\t" << i << "\t" << 4294967294u << "\t" << (float)g << "\t"
         << (float)point_area(g, r[7], r[7], rlim) << "\n";
    }
  }
  return NF_OK;
}

// Node network files: save_coo's node-table and two-column overloads
// (src/conductivity.cu:881-930): "#\tI\tJ\tData" / "#\tI\tJ\tData1\tData2"
// headers, "\tI\tJ\t..." rows.  Electrode rows carry the electrode id
// (4294967295 source, 4294967294 sink) in the first column, as the
// reference's SRC_ELEC / SNK_ELEC node ids.
extern "C" int nf_write_node_coo(nf_handle* h, const char* prefix) {
  if (!h) return fail(nullptr, NF_E_INVALID, "null handle");
  const ContactBuffers& c = h->cb;
  if (!c.nd_valid) return fail(h, NF_E_STATE, "nf_node_network must run before nf_write_node_coo");
  const std::string pre = prefix ? prefix : "";
  const uint64_t nn = c.nd_nodes, ns = c.nd_segments;
  std::vector<uint32_t> rod(nn), sa(ns), sb(ns);
  std::vector<double> t(nn), L(ns), A(ns);
  int rc = nf_get_nodes(h, rod.data(), t.data(), nn);
  if (rc == NF_OK) rc = nf_get_segments(h, sa.data(), sb.data(), L.data(), A.data(), ns);
  uint64_t ne = 0, nel = 0;
  if (rc == NF_OK) rc = nf_get_node_edges(h, nullptr, nullptr, nullptr, 0, &ne);
  if (rc == NF_OK) rc = nf_get_node_electrodes(h, nullptr, nullptr, nullptr, 0, &nel);
  if (rc != NF_OK) return rc;
  std::vector<uint32_t> ea(ne), eb(ne), ek(ne), en(nel);
  std::vector<double> eg(nel);
  std::vector<int32_t> es(nel);
  rc = nf_get_node_edges(h, ea.data(), eb.data(), ek.data(), ne, &ne);
  if (rc == NF_OK) rc = nf_get_node_electrodes(h, en.data(), eg.data(), es.data(), nel, &nel);
  if (rc != NF_OK) return rc;
  const uint32_t n = c.n_rods;
  const uint64_t C = c.n_contacts;
  std::vector<uint32_t> row(n + 1), col(C);
  std::vector<float> gap(C);
  rc = nf_get_contacts(h, row.data(), col.data(), gap.data(), nullptr, C);
  if (rc != NF_OK) return rc;
  std::vector<double> area(C);
  rc = nf_get_contact_area(h, area.data(), C);
  if (rc != NF_OK) return rc;
  std::vector<float> rods((size_t)n * 8);
  uint32_t got = 0;
  rc = nf_get_rods(h, h->post_inst, rods.data(), n, &got);
  if (rc != NF_OK) return rc;
  const double rlim = c.gap_cut;
  std::ofstream fn(pre + "Nodes.tbl.dat"), fsg(pre + "segment.LA.coo"), fi(pre + "internal.DS.coo"),
      fk(pre + "sink.DS.coo"), fs(pre + "source.DS.coo");
  if (!fn.is_open() || !fsg.is_open() || !fi.is_open() || !fk.is_open() || !fs.is_open())
    return fail(h, NF_E_IO, "cannot write node network files");
  fn << "#\tI\tJ\tData\n";
  for (uint64_t g = 0; g < nn; ++g) fn << "\t" << g << "\t" << rod[g] << "\t" << (float)t[g] << "\n";
  fsg << "#\tI\tJ\tData1\tData2\n";
  for (uint64_t k = 0; k < ns; ++k)
    fsg << "\t" << sa[k] << "\t" << sb[k] << "\t" << (float)L[k] << "\t" << (float)A[k] << "\n";
  fi << "#\tI\tJ\tData1\tData2\n";
  for (uint64_t e = 0; e < ne; ++e) {
    const uint32_t k = ek[e];
    const double g = gap[k] > 0.0f ? (double)gap[k] : 0.0;
    fi << "\t" << ea[e] << "\t" << eb[e] << "\t" << (float)g << "\t" << (float)area[k] << "\n";
  }
  fk << "#\tI\tJ\tData1\tData2\n";
  fs << "#\tI\tJ\tData1\tData2\n";
  for (uint64_t e = 0; e < nel; ++e) {
    const double g = eg[e] > 0.0 ? eg[e] : 0.0;
    const double a = rods[8 * (size_t)rod[en[e]] + 7];
    std::ofstream& f = es[e] == 0 ? fs : fk;
    f << "\t" << (es[e] == 0 ? 4294967295u : 4294967294u) << "\t" << en[e] << "\t" << (float)g << "\t"
      << (float)point_area(g, a, a, rlim) << "\n";
  }
  if (!fn.good() || !fsg.good() || !fi.good() || !fk.good() || !fs.good())
    return fail(h, NF_E_IO, "write failed: node network files");
  return NF_OK;
}
