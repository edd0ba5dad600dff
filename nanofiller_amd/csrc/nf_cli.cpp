// nf_cli.cpp - command-line front end with the verbs of the reference's
// main (src/NanoFiller.cpp:48-149):
//
//   nanofiller params.txt generate [write_vtk] [stat_phitheta]
//              [stat_mutangle] [stat_contacts] [expand d] [reduce d]
//              [stat_conduct d [internal]]
//
// d is the field direction as the reference's bit mask: 1 = x, 2 = y, 4 = z
// (init_contacts_map takes 7 - d as the periodic directions,
// src/conductivity.cu:52-58; expand uses axis d >> 1, src/simPrms.cpp:760).
//
// generate          insert/reject generation (cudaMain); prints the count and
//                   the volume fraction reached; with VOLUME_PERCENT_LIST in
//                   the parameter file, state_<vf>.bin milestone dumps
//                   (src/cudaMain.cu:136-142)
// expand d          periodic-copy buffer along direction d (simPrms::expand,
//                   src/simPrms.cpp:744-852): the later verbs see the
//                   expanded packing and box
// write_vtk         output.vtk (write_CNTvtk, src/IO.cpp:38-76)
// stat_phitheta     rho.dat, 32 x 32 phi-theta counts (collect_PhiTheta_distr,
//                   src/statistics_collection.cu:769-861)
// stat_mutangle     gamma.dat, 32 bins of the mutual angle of contacting
//                   pairs (collect_1D_statDistr_weightedFloat, :640-765)
// stat_contacts     contacts per rod: mean and maximum (collect_CntctPerInc)
// stat_vdw          van der Waals energy of the contacting pairs (collect_vdw,
//                   src/statistics_collection.cu:53-154)
// reduce d          dead-end pruning along field direction d (reduce,
//                   src/conductivity.cu:79-138): rods with fewer than 2
//                   contacts are removed round by round; the pruned packing
//                   is the one the later verbs see, and reduced.bin is
//                   written when WRT_BIN is set (src/NanoFiller.cpp:71-90)
// stat_conduct d    contacts, network along direction d, COO files internal.srt.coo
//                   / source.elec.coo / sink.elec.coo (collect_cunduct_data,
//                   src/conductivity.cu:140-423), the DC conductance between
//                   the electrodes (the solve of tools/kirchhoff/kirchhoff.py:
//                   256-272); with "internal" also the node-level files
//                   Nodes.tbl.dat, segment.LA.coo, internal.DS.coo,
//                   sink.DS.coo, source.DS.coo (:213-410)
//
// As in the reference, generate saves output.bin when WRT_BIN is set
// (save_state_BIN, src/stateSR.cpp:54-88), and without generate the rods
// are read from output.bin (read_state_BIN, :90-134).  Verbs outside the
// hot path (plot, stat_nematic, write_ort, write_lmp, stat_exclsurf) are
// reported as unsupported.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "nanofiller.h"

// Axis of a field-direction bit mask (1, 2, 4), or -1.
static int dir_axis(const char* s) {
  const int d = std::atoi(s);
  return d == 1 ? 0 : (d == 2 ? 1 : (d == 4 ? 2 : -1));
}

static int die(nf_handle* h, const char* what, int rc) {
  std::fprintf(stderr, "nanofiller: %s failed (%d): %s\n", what, rc, nf_last_error(h));
  return 1;
}

static bool write_rho(const char* path, int nphi, int ntheta, const std::vector<int64_t>& v) {
  std::ofstream f(path, std::ios::out | std::ios::trunc);
  if (!f.is_open()) return false;
  f << "rho = {";
  for (int j = 0; j < ntheta; ++j) {
    f << "\t{";
    for (int i = 0; i < nphi; ++i) {
      f << std::setw(12) << v[(size_t)i + (size_t)j * nphi];
      if (i != nphi - 1) f << ",";
    }
    f << "}";
    if (j != ntheta - 1) f << "," << std::endl << "\t";
  }
  f << "}";
  return f.good();
}

static bool write_gamma(const char* path, const std::vector<int64_t>& v) {
  std::ofstream f(path, std::ios::out | std::ios::trunc);
  if (!f.is_open() || v.empty()) return false;
  f << "gamma={\t";
  for (size_t i = 0; i + 1 < v.size(); ++i) f << std::setw(10) << v[i] << "," << std::endl;
  f << std::setw(10) << v.back() << "\t}" << std::endl;
  return f.good();
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr,
                 "usage: %s params.txt generate [write_vtk] [stat_phitheta] [stat_mutangle]\n"
                 "       [stat_contacts] [stat_vdw] [expand d] [reduce d] [stat_conduct d [internal]]\n"
                 "       d: field direction 1 (x), 2 (y) or 4 (z)\n",
                 argv[0]);
    return 2;
  }
  nf_params p;
  nf_params_default(&p);
  int rc = nf_params_load(argv[1], &p);
  if (rc != NF_OK) return die(nullptr, "reading parameters", rc);
  if (*nf_last_error(nullptr)) std::printf("note: %s\n", nf_last_error(nullptr));
  rc = nf_params_normalize(&p);
  if (rc != NF_OK) return die(nullptr, "parameters", rc);
  const bool gen = std::strcmp(argv[2], "generate") == 0;
  if (!gen) {
    // the handle must hold the file's rods: size its target from the header
    nf_state_header hd;
    rc = nf_read_state_header("output.bin", &hd);
    if (rc != NF_OK) return die(nullptr, "reading output.bin", rc);
    if (hd.n_rods > p.n_rods) p.n_rods = hd.n_rods;
  }
  nf_handle* h = nullptr;
  rc = nf_create(&p, &h);
  if (rc != NF_OK) return die(nullptr, "device set-up", rc);
  const double PI = 3.14159265358979323846;
  const double V = p.box[0] * p.box[1] * p.box[2];
  uint32_t n = 0;
  nf_gen_stats st;
  std::memset(&st, 0, sizeof(st));
  if (gen) {
    const int nm = nf_params_milestones(argv[1], nullptr, 0);
    if (nm > 0) {
      std::vector<double> ms((size_t)nm);
      nf_params_milestones(argv[1], ms.data(), nm);
      rc = nf_set_milestones(h, ms.data(), nm, ".");
      if (rc != NF_OK) {
        const int r = die(h, "VOLUME_PERCENT_LIST", rc);
        nf_destroy(h);
        return r;
      }
    }
    rc = nf_generate(h, &st);
    if (rc != NF_OK) {
      const int r = die(h, "generate", rc);
      nf_destroy(h);
      return r;
    }
    n = st.n_accepted;
  } else {
    rc = nf_load_state(h, "output.bin", &n);
    if (rc != NF_OK) {
      const int r = die(h, "reading output.bin", rc);
      nf_destroy(h);
      return r;
    }
    std::printf("Simulation state read from file output.bin \n%u inclusions found \n", n);
  }
  std::vector<float> rods((size_t)n * 8 + 8);
  nf_get_rods(h, 0, rods.data(), n, &n);
  if (gen) {
    double vol = 0.0;
    for (uint32_t i = 0; i < n; ++i) vol += PI * (double)rods[8 * i + 7] * rods[8 * i + 7] * rods[8 * i + 6];
    std::printf("%u inclusions generated! \n", n);
    std::printf("Volume fraction of %5.2f percent is achieved! \n", 100.0 * vol / V);
    std::printf("rounds %u, pair tests %llu, rechecks %llu, %.2f ms\n", st.rounds,
                (unsigned long long)st.pair_tests, (unsigned long long)st.rechecks, st.wall_ms);
    if (p.write_state) {
      rc = nf_save_state(h, "output.bin");
      if (rc != NF_OK) {
        const int r = die(h, "writing output.bin", rc);
        nf_destroy(h);
        return r;
      }
      std::printf("Simulation state dumped into file output.bin\n");
    }
  }
  bool have_contacts = false;
  uint64_t nc = 0;
  auto ensure_contacts = [&]() -> int {
    if (have_contacts) return NF_OK;
    const int r = nf_contacts(h, p.contact_gap, &nc);
    if (r == NF_OK) have_contacts = true;
    return r;
  };
  int status = 0;
  for (int ai = gen ? 3 : 2; ai < argc && status == 0; ++ai) {
    const char* v = argv[ai];
    if (!std::strcmp(v, "write_vtk")) {
      std::printf("CNTs created: %u \n", n);
      rc = nf_write_vtk(h, "output.vtk");
      if (rc != NF_OK) status = die(h, "write_vtk", rc);
    } else if (!std::strcmp(v, "stat_phitheta") || !std::strcmp(v, "stat_mutangle")) {
      rc = ensure_contacts();
      if (rc != NF_OK) {
        status = die(h, "contacts", rc);
        break;
      }
      const bool pt = !std::strcmp(v, "stat_phitheta");
      nf_hist_spec spec;
      std::memset(&spec, 0, sizeof(spec));
      spec.bins = 32;
      spec.centre_axis = 0;
      spec.pol_axis[2] = 1.0;
      spec.phi_bins = pt ? 32 : 0;
      spec.theta_bins = pt ? 32 : 0;
      std::vector<int64_t> c(32), po(32), mu(32), ph(32 * 32);
      rc = nf_histograms(h, &spec, c.data(), po.data(), mu.data(), pt ? ph.data() : nullptr);
      if (rc != NF_OK) {
        status = die(h, v, rc);
        break;
      }
      const bool ok = pt ? write_rho("rho.dat", 32, 32, ph) : write_gamma("gamma.dat", mu);
      std::printf("file %s created!\n", pt ? "rho.dat" : "gamma.dat");
      if (!ok) status = die(h, "writing statistics", NF_E_IO);
    } else if (!std::strcmp(v, "stat_contacts")) {
      rc = ensure_contacts();
      if (rc != NF_OK) {
        status = die(h, "contacts", rc);
        break;
      }
      std::vector<uint32_t> row((size_t)n + 1), col(nc ? nc : 1);
      nf_get_contacts(h, row.data(), col.data(), nullptr, nullptr, nc);
      std::vector<uint32_t> per(n, 0);
      for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = row[i]; k < row[i + 1]; ++k) {
          ++per[i];
          ++per[col[k]];
        }
      uint32_t mx = 0;
      for (uint32_t x : per) mx = x > mx ? x : mx;
      std::printf("contacts (gap <= %g um): %llu; per inclusion: mean %.4f, max %u\n", p.contact_gap,
                  (unsigned long long)nc, n ? 2.0 * (double)nc / n : 0.0, mx);
    } else if (!std::strcmp(v, "stat_vdw")) {
      rc = ensure_contacts();
      double ev = 0.0;
      uint64_t np_ = 0;
      if (rc == NF_OK) rc = nf_vdw(h, 15.2, 24100.0, &ev, &np_);
      if (rc != NF_OK) {
        status = die(h, "stat_vdw", rc);
        break;
      }
      std::printf("van der Waals energy over %llu contacting pairs (gap < %g um): %.10e\n",
                  (unsigned long long)np_, p.contact_gap, ev);
    } else if (!std::strcmp(v, "expand")) {
      const int d = ai + 1 < argc ? std::atoi(argv[++ai]) : 0;
      if (dir_axis(std::to_string(d).c_str()) < 0) {
        std::fprintf(stderr, "nanofiller: expand needs a direction 1, 2 or 4\n");
        status = 2;
        break;
      }
      std::printf("Expanding in %d direction:... \n", d);
      nf_handle* g = nullptr;
      rc = nf_expand(h, d, &g);
      if (rc != NF_OK) {
        status = die(h, "expand", rc);
      } else {
        // the expanded packing is the state of the later verbs
        nf_destroy(h);
        h = g;
        nf_get_params(h, 0, &p);
        nf_get_rods(h, 0, nullptr, 0, &n);
        rods.assign((size_t)n * 8 + 8, 0.0f);
        nf_get_rods(h, 0, rods.data(), n, &n);
        have_contacts = false;
        std::printf("New box: %f - %f - %f, new cnt number: %u\n", p.box[0], p.box[1], p.box[2], n);
      }
    } else if (!std::strcmp(v, "reduce")) {
      const int dir = ai + 1 < argc ? dir_axis(argv[++ai]) : -1;
      if (dir < 0) {
        std::fprintf(stderr, "nanofiller: reduce needs a field direction 1, 2 or 4\n");
        status = 2;
        break;
      }
      std::printf("Eliminating inclusions with less then 2 contacts up to %f um: \n", p.contact_gap);
      nf_params q = p;
      q.elec_axis = dir;
      nf_handle* h2 = nullptr;
      rc = nf_create(&q, &h2);
      if (rc == NF_OK) rc = nf_set_rods(h2, 0, rods.data(), n);
      uint64_t nc2 = 0;
      if (rc == NF_OK) rc = nf_contacts(h2, q.contact_gap, &nc2);
      nf_reduce_info ri;
      if (rc == NF_OK) rc = nf_reduce(h2, 2, &ri);
      if (rc == NF_OK) rc = nf_get_rods(h2, 0, rods.data(), n, &n);
      if (rc != NF_OK) {
        status = die(h2, "reduce", rc);
      } else {
        std::printf("Iterations: %u\nInclusions remained: %u \n", ri.rounds, ri.n_after);
        // the pruned packing is the state of the later verbs
        rc = nf_set_rods(h, 0, rods.data(), n);
        have_contacts = false;
        if (rc == NF_OK && p.write_state) rc = nf_save_state(h, "reduced.bin");
        if (rc != NF_OK) status = die(h, "reduce", rc);
      }
      if (h2) nf_destroy(h2);
    } else if (!std::strcmp(v, "stat_conduct")) {
      const int dir = ai + 1 < argc ? dir_axis(argv[++ai]) : -1;
      bool internal = false;
      if (ai + 1 < argc && !std::strcmp(argv[ai + 1], "internal")) {
        internal = true;
        ++ai;
      }
      if (dir < 0) {
        std::fprintf(stderr, "nanofiller: stat_conduct needs a field direction 1, 2 or 4\n");
        status = 2;
        break;
      }
      // the electrode axis is a parameter of the handle: rebuild it with the
      // same rods for this axis
      nf_params q = p;
      q.elec_axis = dir;
      nf_handle* h2 = nullptr;
      rc = nf_create(&q, &h2);
      if (rc == NF_OK) rc = nf_set_rods(h2, 0, rods.data(), n);
      uint64_t nc2 = 0;
      if (rc == NF_OK) rc = nf_contacts(h2, q.contact_gap, &nc2);
      nf_network_info ni;
      if (rc == NF_OK) rc = nf_network(h2, &ni);
      if (rc == NF_OK) rc = nf_write_coo(h2, "");
      if (rc != NF_OK) {
        status = die(h2, "stat_conduct", rc);
      } else {
        std::printf("Building conductivity sparse matrix \n");
        std::printf("contacts %llu, network edges %llu, components %u (largest %u), "
                    "percolating %u, electrode couplings %u / %u, matrix nnz %llu\n",
                    (unsigned long long)nc2, (unsigned long long)ni.n_edges, ni.n_components,
                    ni.largest, ni.percolating, ni.n_source, ni.n_sink,
                    (unsigned long long)ni.nnz);
        if (internal && status == 0) {
          // node level: contact points, segments along the rods, screened
          // contacts removed (the reference's internal-resistance files)
          nf_node_info nd;
          int rc3 = nf_node_network(h2, &nd);
          if (rc3 == NF_OK) rc3 = nf_write_node_coo(h2, "");
          if (rc3 != NF_OK) {
            status = die(h2, "node network", rc3);
          } else {
            std::printf("node network: %u nodes (%u contact and electrode points), %u segments, "
                        "%u contacts kept, %u screened\n",
                        nd.n_nodes, nd.n_points, nd.n_segments, nd.n_edges, nd.n_screened);
          }
        }
        // DC solve between the electrodes (kirchhoff.py's role, real conductances)
        nf_conductance_info ci;
        const int rc2 = nf_conductance(h2, 1e-10, 20000, &ci, nullptr);
        if (rc2 != NF_OK) {
          status = die(h2, "conductance", rc2);
        } else {
          std::printf("effective conductance %.10e (units of TUNNEL_G0) over %u rods, "
                      "%u CG iterations, residual %.2e%s\n",
                      ci.conductance, ci.unknowns, ci.iterations, ci.residual,
                      ci.converged ? "" : " (not converged)");
        }
      }
      if (h2) nf_destroy(h2);
    } else {
      std::fprintf(stderr, "nanofiller: verb %s is not supported by this build\n", v);
      status = 2;
    }
  }
  nf_destroy(h);
  return status;
}
