// nf_nodes.hip - node-level segment network of one packing: contact points
// (rod, t) as nodes, segments along each rod between consecutive nodes, and
// screened contacts removed.
//
// Replaces the internal-resistance branch of collect_cunduct_data
// (src/conductivity.cu:213-410): check_screened_contacts /
// generate_virt_inclusions (conductivity.cu:430-528, inc/conductKernels.h:
// 289-361), get_node_relations (:639-747), get_internal_conduct /
// get_segment_length (:750-836, src/get_contact_data.cu:235-297) and the
// node reduction of src/adjacency_graph.cu:57-219.
//
// Rules (all decisions in FP64 on the FP32 rod data):
//  * a contact (I, J, image) of the network (image not crossing the electrode
//    axis) has contact points at the closest points of the two axes, axial
//    parameters tI, tJ in length units;
//  * it is screened when a third rod's axis comes closer than a_v + a_K to
//    the segment joining the contact points, a_v = min(a_I, a_J) (the
//    virtual inclusion of generate_virt_inclusions);
//  * electrode couplings (the rules of the network matrix) put a point at the
//    rod end facing the plane;
//  * per rod, points are ordered by (t, kind, index) and a point within one
//    radius of its predecessor (t_k - t_{k-1} <= a, L <= L_crit = 1) merges
//    into it; a node's t is the mean of its points;
//  * consecutive nodes of a rod are joined by a segment of length
//    L = (t_b - t_a) / a, cross-section radius A = a.
//
// One wave per contact for the contact points and screening (the lanes
// spread over the rods of the cells whose box lies within reach of the
// contact's midpoint, 64 at a time), one thread per rod for the point
// lists, sorting and merging.
#include "nf_internal.h"

namespace nf {

namespace {

constexpr int NT = 256;

int grid_n(uint64_t n) {
  uint64_t b = (n + NT - 1) / NT;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

__device__ __forceinline__ int img_c(int32_t code, int axis) {
  return (int)(int8_t)((code >> (8 * axis)) & 0xFF);
}

}  // namespace

// Row of every contact of the CSR (one thread per rod).
__global__ void __launch_bounds__(NT) k_nd_rowof(NodeArgs a, uint32_t* __restrict__ rowof) {
  for (uint32_t r = blockIdx.x * NT + threadIdx.x; r < a.n; r += gridDim.x * NT)
    for (uint32_t k = a.rowptr[r]; k < a.rowptr[r + 1]; ++k) rowof[k] = r;
}

// Contact points and screening, one wave per contact; the lanes look up 64
// cells around the contact's virtual inclusion at once, and the rods of the
// cells within reach are packed into 64-lane batches (a lane finds its cell
// by a binary search of the cells' prefix counts, held one per lane and read
// by shuffles), so every batch's loads are in flight together.  All loop
// bounds are wave-uniform.
__global__ void __launch_bounds__(NT) k_nd_points(NodeArgs a) {
  const uint32_t nc = a.rowptr[a.n];
  const uint32_t lane = threadIdx.x & (WAVE - 1);
  const uint32_t w0 = (blockIdx.x * NT + threadIdx.x) / WAVE, nw = gridDim.x * NT / WAVE;
  // one wave per contact: every lane holds the same contact values
  for (uint32_t k = w0; k < nc; k += nw) {
    const uint32_t r = a.rowof[k];
    const float4 PI = a.P[r], DI = a.D[r];
    {
      const int32_t code = a.img[k];
      if (img_c(code, a.axis) != 0) {
        if (lane == 0) a.pflag[k] = NODE_CROSSING;
        continue;
      }
      const uint32_t j = a.J[k];
      const float4 PJ = a.P[j], DJ = a.D[j];
      const double ux = (double)img_c(code, 0) * a.box[0], uy = (double)img_c(code, 1) * a.box[1],
                   uz = (double)img_c(code, 2) * a.box[2];
      double s = 0.0, t = 0.0;
      bool w;
      (void)pair_closest_f64(PI, DI, PJ, DJ, ux, uy, uz, 0.0, &w, &s, &t);
      const double tI = s * (double)PI.w, tJ = t * (double)PJ.w;
      if (lane == 0) {
        a.ptI[k] = tI;
        a.ptJ[k] = tJ;
      }
      // the segment between the contact points
      const double ax = (double)PI.x + tI * (double)DI.x, ay = (double)PI.y + tI * (double)DI.y,
                   az = (double)PI.z + tI * (double)DI.z;
      const double bx = ((double)PJ.x + ux) + tJ * (double)DJ.x,
                   by = ((double)PJ.y + uy) + tJ * (double)DJ.y,
                   bz = ((double)PJ.z + uz) + tJ * (double)DJ.z;
      const double av = (double)DI.w < (double)DJ.w ? (double)DI.w : (double)DJ.w;
      const double mx = 0.5 * (ax + bx), my = 0.5 * (ay + by), mz = 0.5 * (az + bz);
      const double hx = 0.5 * (bx - ax), hy = 0.5 * (by - ay), hz = 0.5 * (bz - az);
      const double hs = sqrt_cr((hx * hx + hy * hy) + hz * hz);
      // cells around the midpoint that can hold a screening rod
      const double reach = a.hl_max + hs + av + a.a_max;
      int c[3], kc[3], span[3];
      const double mp[3] = {mx, my, mz};
      for (int d = 0; d < 3; ++d) {
        double v = mp[d] - floor(mp[d] / a.box[d]) * a.box[d];
        int ci = (int)(v * a.cell_scale[d]);
        ci = ci < 0 ? 0 : (ci >= a.ncell[d] ? a.ncell[d] - 1 : ci);
        c[d] = ci;
        kc[d] = (int)floor(reach * a.cell_scale[d]) + 1;
        span[d] = 2 * kc[d] + 1;
      }
      // the midpoint's own periodic image relative to the primary box
      const double base[3] = {floor(mx / a.box[0]) * a.box[0], floor(my / a.box[1]) * a.box[1],
                              floor(mz / a.box[2]) * a.box[2]};
      // FP32 sieve radius ahead of the FP64 test
      const float fr = (float)(hs + av);
      const uint32_t ncand = (uint32_t)(span[0] * span[1] * span[2]);
      bool screened = false;
      for (uint32_t b0 = 0; b0 < ncand && !screened; b0 += WAVE) {
        // ---- this lane's cell of the block: first rod, rod count, image
        const uint32_t o = b0 + lane;
        uint32_t cnt = 0u, st0 = 0u, imc = 0u;
        if (o < ncand) {
          const int off[3] = {(int)(o % (uint32_t)span[0]) - kc[0],
                              (int)((o / (uint32_t)span[0]) % (uint32_t)span[1]) - kc[1],
                              (int)(o / (uint32_t)(span[0] * span[1])) - kc[2]};
          int m[3], im[3];
          double gap2 = 0.0;  // squared distance from the midpoint to the cell's box
          for (int d = 0; d < 3; ++d) {
            const int raw = c[d] + off[d];
            int mm = raw % a.ncell[d];
            if (mm < 0) mm += a.ncell[d];
            im[d] = (raw - mm) / a.ncell[d];
            m[d] = mm;
            const double e = a.box[d] / (double)a.ncell[d];
            const double lo = (double)raw * e + base[d], hi = lo + e;
            const double dd = mp[d] < lo ? lo - mp[d] : (mp[d] > hi ? mp[d] - hi : 0.0);
            gap2 += dd * dd;
          }
          // rods are binned by centre: none in a cell farther than reach
          if (gap2 <= reach * reach * 1.000001) {
            const uint32_t cell = (uint32_t)(m[0] + a.ncell[0] * (m[1] + a.ncell[1] * m[2]));
            st0 = a.cstart[cell];
            cnt = a.ccount[cell];
            imc = (uint32_t)(im[0] & 0xFF) | ((uint32_t)(im[1] & 0xFF) << 8) | ((uint32_t)(im[2] & 0xFF) << 16);
          }
        }
        // the block's cells back to back: exclusive prefix of the counts
        uint32_t incl = cnt;
#pragma unroll
        for (int sft = 1; sft < WAVE; sft <<= 1) {
          const uint32_t tv = (uint32_t)__shfl_up((int)incl, sft, WAVE);
          incl += lane >= (uint32_t)sft ? tv : 0u;
        }
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, WAVE - 1);
        const uint32_t pre = incl - cnt;
        for (uint32_t t0 = 0; t0 < total && !screened; t0 += WAVE) {
          const uint32_t pos = t0 + lane;
          // the last lane whose prefix is <= pos holds rod pos (a lane with
          // no rods shares the next lane's prefix)
          uint32_t kk = 0u;
#pragma unroll
          for (uint32_t step = WAVE / 2; step > 0; step >>= 1) {
            const uint32_t pv = (uint32_t)__shfl((int)pre, (int)(kk + step), WAVE);
            if (pv <= pos) kk += step;
          }
          const uint32_t cst = (uint32_t)__shfl((int)st0, (int)kk, WAVE);
          const uint32_t cpre = (uint32_t)__shfl((int)pre, (int)kk, WAVE);
          const uint32_t cim = (uint32_t)__shfl((int)imc, (int)kk, WAVE);
          bool hit = false;
          if (pos < total) {
            const uint32_t q = cst + (pos - cpre);
            const double sh[3] = {base[0] + (double)img_c((int32_t)cim, 0) * a.box[0],
                                  base[1] + (double)img_c((int32_t)cim, 1) * a.box[1],
                                  base[2] + (double)img_c((int32_t)cim, 2) * a.box[2]};
            const uint32_t kq = a.cid[q];
            const float4 PK = a.cP[q], DK = a.cD[q];
            // FP32 sieve: the midpoint in the cell's frame, rounded once
            // (error far below the 1e-4 slack)
            const float fx = (float)(mx - sh[0]), fy = (float)(my - sh[1]), fz = (float)(mz - sh[2]);
            const float gx = PK.x - fx, gy = PK.y - fy, gz = PK.z - fz;
            const float gl = (PK.w + fr + DK.w) * 1.0001f + 1e-4f;
            const bool near = fmaf(gz, gz, fmaf(gy, gy, gx * gx)) <= gl * gl;
            if (near && kq != r && kq != j) {
              const double R = av + (double)DK.w;
              // conservative centre test, then the exact segment distance
              const double cx = ((double)PK.x + sh[0]) - mx, cy = ((double)PK.y + sh[1]) - my,
                           cz = ((double)PK.z + sh[2]) - mz;
              const double lim = ((double)PK.w + hs + R) * 1.000000001;
              if ((cx * cx + cy * cy) + cz * cz <= lim * lim) {
                const double d2 = seg_seg_d2(mx, my, mz, hx, hy, hz, (double)PK.x + sh[0],
                                             (double)PK.y + sh[1], (double)PK.z + sh[2],
                                             (double)PK.w * (double)DK.x, (double)PK.w * (double)DK.y,
                                             (double)PK.w * (double)DK.z);
                hit = d2 < R * R;
              }
            }
          }
          screened = __ballot(hit) != 0ull;
        }
      }
      if (lane == 0) {
        a.pflag[k] = screened ? NODE_SCREENED : NODE_KEPT;
        if (!screened) {
          atomicAdd(&a.npt[r], 1u);
          atomicAdd(&a.npt[j], 1u);
        }
      }
    }
  }
}

// Electrode couplings (the network's rule, k_net_elec) and their points.
__global__ void __launch_bounds__(NT) k_nd_elec(NodeArgs a) {
  for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < a.n; i += gridDim.x * NT) {
    const float4 p = a.P[i], d = a.D[i];
    const float pa = a.axis == 0 ? p.x : (a.axis == 1 ? p.y : p.z);
    const float ca = a.axis == 0 ? d.x : (a.axis == 1 ? d.y : d.z);
    const double ext = fabs((double)p.w * (double)ca);
    const double pos = (double)pa, rad = (double)d.w;
    const double gs = (pos - ext) - rad;
    const double gk = (a.L - (pos + ext)) - rad;
    const bool cs = gs <= a.gcut, ck = gk <= a.gcut;
    a.esrc[i] = cs ? 1u : 0u;
    a.esnk[i] = ck ? 1u : 0u;
    a.egs[i] = gs;
    a.egk[i] = gk;
    a.npt[i] += (cs ? 1u : 0u) + (ck ? 1u : 0u);
  }
}

// Points into per-rod lists (unordered; sorted next).
__global__ void __launch_bounds__(NT) k_nd_scatter(NodeArgs a) {
  for (uint32_t r = blockIdx.x * NT + threadIdx.x; r < a.n; r += gridDim.x * NT) {
    for (uint32_t k = a.rowptr[r]; k < a.rowptr[r + 1]; ++k) {
      if (a.pflag[k] != NODE_KEPT) continue;
      const uint32_t j = a.J[k];
      const uint32_t p0 = a.pstart[r] + atomicAdd(&a.pfill[r], 1u);
      a.pt_t[p0] = a.ptI[k];
      a.pt_key[p0] = ((unsigned long long)PT_CONTACT_I << 32) | k;
      const uint32_t p1 = a.pstart[j] + atomicAdd(&a.pfill[j], 1u);
      a.pt_t[p1] = a.ptJ[k];
      a.pt_key[p1] = ((unsigned long long)PT_CONTACT_J << 32) | k;
    }
    // electrode points: the end facing each plane
    const float4 d = a.D[r];
    const float ca = a.axis == 0 ? d.x : (a.axis == 1 ? d.y : d.z);
    const double hl = (double)a.P[r].w;
    if (a.esrc[r]) {
      const uint32_t p = a.pstart[r] + atomicAdd(&a.pfill[r], 1u);
      a.pt_t[p] = ca >= 0.0f ? -hl : hl;
      a.pt_key[p] = ((unsigned long long)PT_SOURCE << 32) | r;
    }
    if (a.esnk[r]) {
      const uint32_t p = a.pstart[r] + atomicAdd(&a.pfill[r], 1u);
      a.pt_t[p] = ca >= 0.0f ? hl : -hl;
      a.pt_key[p] = ((unsigned long long)PT_SINK << 32) | r;
    }
  }
}

// Per rod: sort the points by (t, kind, index), merge chains closer than
// one radius, count nodes and segments.
__global__ void __launch_bounds__(NT) k_nd_sort(NodeArgs a) {
  for (uint32_t r = blockIdx.x * NT + threadIdx.x; r < a.n; r += gridDim.x * NT) {
    const uint32_t b = a.pstart[r], e = a.pstart[r + 1];
    for (uint32_t i = b + 1; i < e; ++i) {
      const double t = a.pt_t[i];
      const unsigned long long kk = a.pt_key[i];
      uint32_t q = i;
      while (q > b && (a.pt_t[q - 1] > t || (a.pt_t[q - 1] == t && a.pt_key[q - 1] > kk))) {
        a.pt_t[q] = a.pt_t[q - 1];
        a.pt_key[q] = a.pt_key[q - 1];
        --q;
      }
      a.pt_t[q] = t;
      a.pt_key[q] = kk;
    }
    const double rad = (double)a.D[r].w;
    uint32_t g = 0;
    for (uint32_t i = b; i < e; ++i) {
      if (i > b && !(a.pt_t[i] - a.pt_t[i - 1] <= rad)) ++g;
      a.pt_grp[i] = g;
    }
    a.ngrp[r] = e > b ? g + 1 : 0u;
    a.nseg[r] = e > b ? g : 0u;
  }
}

// Nodes (mean t of their points), segments, and the node of every point.
__global__ void __launch_bounds__(NT) k_nd_nodes(NodeArgs a) {
  for (uint32_t r = blockIdx.x * NT + threadIdx.x; r < a.n; r += gridDim.x * NT) {
    const uint32_t b = a.pstart[r], e = a.pstart[r + 1];
    if (e == b) continue;
    const uint32_t g0 = a.gstart[r];
    const double rad = (double)a.D[r].w;
    double sum = 0.0, prev_t = 0.0;
    uint32_t cnt = 0, g = 0;
    for (uint32_t i = b; i <= e; ++i) {
      if (i == e || a.pt_grp[i] != g) {
        const double tm = sum / (double)cnt;
        a.node_rod[g0 + g] = r;
        a.node_t[g0 + g] = tm;
        if (g > 0) {
          const uint32_t sgi = a.sstart[r] + g - 1;
          a.seg_a[sgi] = g0 + g - 1;
          a.seg_b[sgi] = g0 + g;
          a.seg_L[sgi] = (tm - prev_t) / rad;
          a.seg_A[sgi] = rad;
        }
        prev_t = tm;
        if (i == e) break;
        g = a.pt_grp[i];
        sum = 0.0;
        cnt = 0;
      }
      sum += a.pt_t[i];
      ++cnt;
      const unsigned long long key = a.pt_key[i];
      const uint32_t kind = (uint32_t)(key >> 32), idx = (uint32_t)key;
      const uint32_t node = g0 + a.pt_grp[i];
      if (kind == PT_CONTACT_I)
        a.edge_a[idx] = node;
      else if (kind == PT_CONTACT_J)
        a.edge_b[idx] = node;
      else if (kind == PT_SOURCE)
        a.src_node[idx] = node;
      else
        a.snk_node[idx] = node;
    }
  }
}

// Counts for the info record: kept and screened contacts, electrode points
// (wave sums, one atomic per wave).
__global__ void __launch_bounds__(NT) k_nd_count(NodeArgs a, uint32_t nc,
                                                 unsigned long long* __restrict__ out) {
  unsigned long long kept = 0, scr = 0, src = 0, snk = 0;
  const uint32_t m = nc > a.n ? nc : a.n;
  for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < m; i += gridDim.x * NT) {
    if (i < nc) {
      kept += a.pflag[i] == NODE_KEPT ? 1u : 0u;
      scr += a.pflag[i] == NODE_SCREENED ? 1u : 0u;
    }
    if (i < a.n) {
      src += a.esrc[i];
      snk += a.esnk[i];
    }
  }
  kept = wave_sum_u64(kept);
  scr = wave_sum_u64(scr);
  src = wave_sum_u64(src);
  snk = wave_sum_u64(snk);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (kept) atomicAdd(&out[0], kept);
    if (scr) atomicAdd(&out[1], scr);
    if (src) atomicAdd(&out[2], src);
    if (snk) atomicAdd(&out[3], snk);
  }
}

// Compact tables for the getters, so that they copy only what they return:
// the kept contacts' edges in contact order, and the electrode points,
// source side then sink side, each in rod order.  Ranks come from scans of
// the flags.
__global__ void __launch_bounds__(NT) k_nd_kept_flag(const uint32_t* __restrict__ pflag, uint32_t nc,
                                                     uint32_t* __restrict__ f) {
  for (uint32_t k = blockIdx.x * NT + threadIdx.x; k < nc; k += gridDim.x * NT) f[k] = pflag[k] == NODE_KEPT ? 1u : 0u;
}

__global__ void __launch_bounds__(NT) k_nd_edges_out(NodeArgs a, uint32_t nc, const uint32_t* __restrict__ f,
                                                     const uint32_t* __restrict__ rank, uint32_t* __restrict__ oa,
                                                     uint32_t* __restrict__ ob, uint32_t* __restrict__ ok) {
  for (uint32_t k = blockIdx.x * NT + threadIdx.x; k < nc; k += gridDim.x * NT) {
    if (!f[k]) continue;
    const uint32_t j = rank[k];
    oa[j] = a.edge_a[k];
    ob[j] = a.edge_b[k];
    ok[j] = k;
  }
}

__global__ void __launch_bounds__(NT) k_nd_elec_out(NodeArgs a, const uint32_t* __restrict__ rank_src,
                                                    const uint32_t* __restrict__ rank_snk, uint32_t n_src,
                                                    uint32_t* __restrict__ onode, double* __restrict__ ogap,
                                                    int32_t* __restrict__ oside) {
  for (uint32_t i = blockIdx.x * NT + threadIdx.x; i < a.n; i += gridDim.x * NT) {
    if (a.esrc[i]) {
      const uint32_t j = rank_src[i];
      onode[j] = a.src_node[i];
      ogap[j] = a.egs[i];
      oside[j] = 0;
    }
    if (a.esnk[i]) {
      const uint32_t j = n_src + rank_snk[i];
      onode[j] = a.snk_node[i];
      ogap[j] = a.egk[i];
      oside[j] = 1;
    }
  }
}

void launch_node_kept_flag(const uint32_t* pflag, uint32_t nc, uint32_t* f, hipStream_t s) {
  if (nc) hipLaunchKernelGGL(k_nd_kept_flag, dim3(grid_n(nc)), dim3(NT), 0, s, pflag, nc, f);
}

void launch_node_tables(const NodeArgs& a, uint32_t nc, const uint32_t* kept, const uint32_t* rank_kept,
                        uint32_t* oa, uint32_t* ob, uint32_t* ok, const uint32_t* rank_src,
                        const uint32_t* rank_snk, uint32_t n_src, uint32_t* onode, double* ogap, int32_t* oside,
                        hipStream_t s) {
  if (nc) hipLaunchKernelGGL(k_nd_edges_out, dim3(grid_n(nc)), dim3(NT), 0, s, a, nc, kept, rank_kept, oa, ob, ok);
  if (a.n)
    hipLaunchKernelGGL(k_nd_elec_out, dim3(grid_n(a.n)), dim3(NT), 0, s, a, rank_src, rank_snk, n_src, onode, ogap,
                       oside);
}

void launch_node_count(const NodeArgs& a, uint32_t nc, unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(k_nd_count, dim3(grid_n(nc > a.n ? nc : a.n)), dim3(NT), 0, s, a, nc, out);
}

void launch_node_points(const NodeArgs& a, uint32_t* rowof, hipStream_t s) {
  if (a.n) hipLaunchKernelGGL(k_nd_rowof, dim3(grid_n(a.n)), dim3(NT), 0, s, a, rowof);
  hipLaunchKernelGGL(k_nd_points, dim3(8192), dim3(NT), 0, s, a);  // waves stride over the contacts
  hipLaunchKernelGGL(k_nd_elec, dim3(grid_n(a.n)), dim3(NT), 0, s, a);
}

void launch_node_lists(const NodeArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_nd_scatter, dim3(grid_n(a.n)), dim3(NT), 0, s, a);
  hipLaunchKernelGGL(k_nd_sort, dim3(grid_n(a.n)), dim3(NT), 0, s, a);
}

void launch_node_nodes(const NodeArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_nd_nodes, dim3(grid_n(a.n)), dim3(NT), 0, s, a);
}

}  // namespace nf
