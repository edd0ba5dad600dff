"""ctypes binding of the FP64 CPU oracle (oracle/nf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, as the checker.  The shipped library never loads
it.  Build with ``make -C oracle``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnforacle.so")
_lib = None


class GenStats(C.Structure):
    _fields_ = [
        ("n_accepted", C.c_uint32),
        ("rounds", C.c_uint32),
        ("status", C.c_int32),
        ("pad", C.c_uint32),
        ("candidates", C.c_uint64),
        ("proposals", C.c_uint64),
        ("pair_tests", C.c_uint64),
        ("rejected_old", C.c_uint64),
        ("rejected_new", C.c_uint64),
    ]

    def to_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad"}


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        from nanofiller_amd import HistSpec, NetworkInfo, Params

        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        sigs = [
            ("orc_philox4x32_10", None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
            ("orc_u01", C.c_double, [C.c_uint32, C.c_uint32]),
            ("orc_exp", C.c_double, [C.c_double]),
            ("orc_log", C.c_double, [C.c_double]),
            ("orc_agg_centres", None, [P(Params), C.c_void_p]),
            ("orc_vmf_basis", None, [C.c_void_p, C.c_void_p, C.c_void_p]),
            ("orc_candidate", C.c_int, [P(Params), C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
            ("orc_mesh_density", C.c_double, [P(Params), C.c_double, C.c_double, C.c_double]),
            ("orc_node_points", None, [P(Params), C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int]),
            ("orc_pair_d2", C.c_double, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, P(C.c_int)]),
            ("orc_contact_area", C.c_double, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
            ("orc_contact_areas", None, [P(Params), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_int64, C.c_double, C.c_void_p]),
            ("orc_pair_batch", None, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                      C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
            ("orc_generate", C.c_int, [P(Params), C.c_void_p, P(GenStats), C.c_void_p, C.c_int]),
            ("orc_contacts", C.c_int64, [P(Params), C.c_void_p, C.c_uint32, C.c_double, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                         P(C.c_uint64), C.c_int]),
            ("orc_histograms", C.c_int, [P(Params), C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                         C.c_int64, P(HistSpec), C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]),
            ("orc_atan", C.c_double, [C.c_double]),
            ("orc_expand", C.c_int64, [C.c_void_p, C.c_uint32, C.c_int, C.c_double, C.c_double,
                                       C.c_void_p, C.c_int64]),
            ("orc_reduce", C.c_uint32, [P(Params), C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_int64, C.c_double, C.c_int32, C.c_void_p,
                                        P(C.c_uint32)]),
            ("orc_vdw_pair", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                       C.c_double, P(C.c_double)]),
            ("orc_vdw", C.c_uint64, [P(Params), C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_int64, C.c_double, C.c_double, C.c_double,
                                     P(C.c_double)]),
            ("orc_centre_distance", C.c_uint64, [P(Params), C.c_void_p, C.c_uint32, C.c_int32,
                                                 C.c_double, C.c_void_p, P(C.c_uint64),
                                                 P(C.c_uint64)]),
            ("orc_network", C.c_int64, [P(Params), C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int64, P(NetworkInfo)]),
        ]
        for name, res, args in sigs:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def philox(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def pair_d2(ri, rj, shift, R) -> tuple[float, bool]:
    a = np.ascontiguousarray(ri, np.float32)
    b = np.ascontiguousarray(rj, np.float32)
    s = np.ascontiguousarray(shift, np.float64)
    w = C.c_int()
    d2 = lib().orc_pair_d2(a.ctypes.data, b.ctypes.data, s.ctypes.data, float(R), C.byref(w))
    return d2, bool(w.value)


def contact_area(ri, rj, shift, rlim) -> float:
    """Contact area S of rods ri, rj (8 floats each), rj displaced by shift."""
    a = np.ascontiguousarray(ri, np.float32)
    b = np.ascontiguousarray(rj, np.float32)
    s = np.ascontiguousarray(shift, np.float64)
    return lib().orc_contact_area(a.ctypes.data, b.ctypes.data, s.ctypes.data, float(rlim))


def contact_areas(p, rods, I, J, image, rlim) -> np.ndarray:
    r = np.ascontiguousarray(rods, np.float32)
    I = np.ascontiguousarray(I, np.uint32)
    J = np.ascontiguousarray(J, np.uint32)
    im = np.ascontiguousarray(image, np.int32)
    out = np.zeros(len(I), np.float64)
    lib().orc_contact_areas(C.byref(p), r.ctypes.data, I.ctypes.data, J.ctypes.data, im.ctypes.data,
                            len(I), float(rlim), out.ctypes.data)
    return out


def pair_batch(ra, rb, shifts, contact: bool, sep_or_gap: float, soft_frac: float = 1.0):
    ra = np.ascontiguousarray(ra, np.float32)
    rb = np.ascontiguousarray(rb, np.float32)
    sh = np.ascontiguousarray(shifts, np.float64)
    n = ra.shape[0]
    d2 = np.zeros(n, np.float64)
    w = np.zeros(n, np.int32)
    cut = np.zeros(n, np.int32)
    lib().orc_pair_batch(n, ra.ctypes.data, rb.ctypes.data, sh.ctypes.data, int(contact),
                         float(sep_or_gap), float(soft_frac), d2.ctypes.data, w.ctypes.data,
                         cut.ctypes.data)
    return d2, w.astype(bool), cut.astype(bool)


def mesh_density(p, x: float, y: float, z: float) -> float:
    """Density map value at a point (trilinear, periodic; params' map)."""
    return lib().orc_mesh_density(C.byref(p), x, y, z)


def candidate(p, rnd: int, slot: int, agg=None):
    out = np.zeros(8, np.float32)
    a = np.ascontiguousarray(agg if agg is not None else np.zeros(3), np.float64)
    v = lib().orc_candidate(C.byref(p), a.ctypes.data, rnd, slot, out.ctypes.data)
    return v, out


def agg_centres(p) -> np.ndarray:
    out = np.zeros((max(p.n_agg, 1), 3), np.float64)
    lib().orc_agg_centres(C.byref(p), out.ctypes.data)
    return out[: p.n_agg]


def generate(p, threads: int = 1):
    rods = np.zeros((p.n_rods, 8), np.float32)
    st = GenStats()
    acc = np.zeros(max(p.max_rounds, 1), np.uint32)
    rc = lib().orc_generate(C.byref(p), rods.ctypes.data, C.byref(st), acc.ctypes.data, threads)
    if rc != 0:
        raise RuntimeError(f"oracle generate failed {rc}")
    return rods[: st.n_accepted], st.to_dict(), acc[: st.rounds]


def contacts(p, rods: np.ndarray, gap: float, threads: int = 1):
    r = np.ascontiguousarray(rods, np.float32)
    n = r.shape[0]
    tests = C.c_uint64()
    cnt = lib().orc_contacts(C.byref(p), r.ctypes.data, n, gap, None, None, None, None, None, 0,
                             C.byref(tests), threads)
    I = np.zeros(cnt, np.uint32)
    J = np.zeros(cnt, np.uint32)
    g32 = np.zeros(cnt, np.float32)
    g64 = np.zeros(cnt, np.float64)
    im = np.zeros(cnt, np.int32)
    cnt2 = lib().orc_contacts(C.byref(p), r.ctypes.data, n, gap, I.ctypes.data, J.ctypes.data,
                              g32.ctypes.data, g64.ctypes.data, im.ctypes.data, cnt, C.byref(tests),
                              threads)
    assert cnt2 == cnt
    return {"I": I, "J": J, "gap": g32, "gap64": g64, "image": im, "pair_tests": tests.value}


def histograms(p, rods, I, J, bins=64, centre_axis=0, pol_axis=(1.0, 0.0, 0.0), phi_bins=0,
               theta_bins=0):
    from nanofiller_amd import HistSpec

    spec = HistSpec()
    spec.bins = bins
    spec.centre_axis = centre_axis
    for i in range(3):
        spec.pol_axis[i] = pol_axis[i]
    spec.phi_bins = phi_bins
    spec.theta_bins = theta_bins
    r = np.ascontiguousarray(rods, np.float32)
    I = np.ascontiguousarray(I, np.uint32)
    J = np.ascontiguousarray(J, np.uint32)
    c = np.zeros(bins, np.int64)
    pp = np.zeros(bins, np.int64)
    m = np.zeros(bins, np.int64)
    pt = np.zeros(max(phi_bins * theta_bins, 1), np.int64)
    rc = lib().orc_histograms(C.byref(p), r.ctypes.data, r.shape[0], I.ctypes.data, J.ctypes.data,
                              len(I), C.byref(spec), c.ctypes.data, pp.ctypes.data, m.ctypes.data,
                              pt.ctypes.data if phi_bins else None)
    if rc != 0:
        raise RuntimeError("oracle histograms failed")
    out = {"centre": c, "polar": pp, "mutual": m}
    if phi_bins:
        out["phitheta"] = pt.reshape(theta_bins, phi_bins)
    return out


def network(p, rods, I, J, gap64, image):
    from nanofiller_amd import NetworkInfo

    r = np.ascontiguousarray(rods, np.float32)
    n = r.shape[0]
    I = np.ascontiguousarray(I, np.uint32)
    J = np.ascontiguousarray(J, np.uint32)
    g = np.ascontiguousarray(gap64, np.float64)
    im = np.ascontiguousarray(image, np.int32)
    info = NetworkInfo()
    nnz = lib().orc_network(C.byref(p), r.ctypes.data, n, I.ctypes.data, J.ctypes.data, g.ctypes.data,
                            im.ctypes.data, len(I), None, None, None, None, 0, C.byref(info))
    labels = np.zeros(n, np.uint32)
    row = np.zeros(n + 3, np.uint32)
    col = np.zeros(nnz, np.uint32)
    val = np.zeros(nnz, np.float64)
    lib().orc_network(C.byref(p), r.ctypes.data, n, I.ctypes.data, J.ctypes.data, g.ctypes.data,
                      im.ctypes.data, len(I), labels.ctypes.data, row.ctypes.data, col.ctypes.data,
                      val.ctypes.data, nnz, C.byref(info))
    return info.to_dict(), {"labels": labels, "row_ptr": row, "col": col, "val": val}


def expand(rods, axis, L, B):
    """Periodic copies of the rods within B of the two faces normal to
    `axis`, every coordinate along it shifted by B (orc_expand)."""
    r = np.ascontiguousarray(rods, np.float32)
    n = lib().orc_expand(r.ctypes.data, r.shape[0], int(axis), float(L), float(B), None, 0)
    out = np.zeros((max(n, 1), 8), np.float32)
    lib().orc_expand(r.ctypes.data, r.shape[0], int(axis), float(L), float(B), out.ctypes.data, n)
    return out[:n]


def reduce(p, rods, I, J, image, gcut, min_cont=2):
    """Dead-end pruning of a contact list: (keep mask, rounds)."""
    r = np.ascontiguousarray(rods, np.float32)
    I = np.ascontiguousarray(I, np.uint32)
    J = np.ascontiguousarray(J, np.uint32)
    im = np.ascontiguousarray(image, np.int32)
    keep = np.zeros(max(r.shape[0], 1), np.uint8)
    rounds = C.c_uint32()
    lib().orc_reduce(C.byref(p), r.ctypes.data, r.shape[0], I.ctypes.data, J.ctypes.data,
                     im.ctypes.data, len(I), float(gcut), int(min_cont), keep.ctypes.data,
                     C.byref(rounds))
    return keep[: r.shape[0]].astype(bool), rounds.value


def atan(x: float) -> float:
    return lib().orc_atan(float(x))


def vdw_pair(r1, r2, shift, dlim, A=15.2, B=24100.0):
    a = np.ascontiguousarray(r1, np.float32)
    b = np.ascontiguousarray(r2, np.float32)
    sh = np.ascontiguousarray(shift, np.float64)
    e = C.c_double()
    ok = lib().orc_vdw_pair(a.ctypes.data, b.ctypes.data, sh.ctypes.data, dlim, A, B, C.byref(e))
    return (e.value if ok else None)


def vdw(p, rods, I, J, image, dlim, A=15.2, B=24100.0):
    r = np.ascontiguousarray(rods, np.float32)
    I = np.ascontiguousarray(I, np.uint32)
    J = np.ascontiguousarray(J, np.uint32)
    im = np.ascontiguousarray(image, np.int32)
    e = C.c_double()
    n = lib().orc_vdw(C.byref(p), r.ctypes.data, r.shape[0], I.ctypes.data, J.ctypes.data,
                      im.ctypes.data, len(I), dlim, A, B, C.byref(e))
    return e.value, n


def centre_distance(p, rods, bins, width):
    r = np.ascontiguousarray(rods, np.float32)
    counts = np.zeros(bins, np.int64)
    s1, s2 = C.c_uint64(), C.c_uint64()
    n = lib().orc_centre_distance(C.byref(p), r.ctypes.data, r.shape[0], bins, width,
                                  counts.ctypes.data, C.byref(s1), C.byref(s2))
    return {"counts": counts, "pairs": n, "sum_d_q24": s1.value, "sum_d2_q16": s2.value}


def conductance(net: dict):
    """DC conductance between the electrodes of a Kirchhoff network (source at
    potential 1, sink at 0), restating tools/kirchhoff/kirchhoff.py:256-272
    (solver_type 0: scipy spsolve of the Laplacian with the source couplings
    as right-hand side, current I = sum (1 - V_i) B_i) for real conductances.

    Only components touching both electrodes are solved; rods of components
    touching only the source / only the sink sit at 1 / 0, floating ones are
    NaN.  Returns the source and sink currents, the unknown count and every
    rod's potential.
    """
    import scipy.sparse as sp
    from scipy.sparse.linalg import spsolve

    labels = np.asarray(net["labels"], np.int64)
    n = len(labels)
    row = np.asarray(net["row_ptr"], np.int64)
    m = sp.csr_matrix((net["val"], net["col"], row), shape=(n + 2, n + 2))
    flag = np.zeros(n, np.int64)
    for e, bit in ((n, 1), (n + 1, 2)):
        cols = net["col"][row[e]:row[e + 1]]
        for j in cols[cols < n]:
            flag[labels[j]] |= bit
    f = flag[labels] if n else flag
    act = np.flatnonzero(f == 3)
    v = np.full(n, np.nan)
    v[f == 1] = 1.0
    v[f == 2] = 0.0
    isrc = isnk = 0.0
    if len(act):
        a = m[act][:, act].tocsc()
        b = -m[act][:, [n]].toarray().ravel()
        bk = -m[act][:, [n + 1]].toarray().ravel()
        x = spsolve(a, b) if len(act) > 1 else np.array([b[0] / a.toarray()[0, 0]])
        v[act] = x
        isrc = float(np.dot(b, 1.0 - x))
        isnk = float(np.dot(bk, x))
    return {"conductance": isrc, "current_sink": isnk, "unknowns": len(act), "potentials": v}


def node_network(p, rods, I, J, image, threads: int = 8):
    """Node-level segment network of a contact list (the internal-resistance
    branch of collect_cunduct_data, src/conductivity.cu:213-410): contact
    points and screening from orc_node_points, electrode points at the rod
    ends facing the planes (the couplings of the network matrix), per-rod
    points ordered by (t, kind, index), chains of points within one radius
    merged (remove_redundant_nodes / reduce_node_tbl, src/adjacency_graph.cu:
    57-155; node t = mean of its points), segments L = dt / a between
    consecutive nodes (get_segment_length / recalculate_segments)."""
    r = np.ascontiguousarray(rods, np.float32)
    n = r.shape[0]
    I = np.ascontiguousarray(I, np.uint32)
    J = np.ascontiguousarray(J, np.uint32)
    im = np.ascontiguousarray(image, np.int32)
    nc = len(I)
    tI = np.zeros(max(nc, 1))
    tJ = np.zeros(max(nc, 1))
    flag = np.zeros(max(nc, 1), np.uint32)
    lib().orc_node_points(C.byref(p), r.ctypes.data, n, I.ctypes.data, J.ctypes.data, im.ctypes.data,
                          nc, tI.ctypes.data, tJ.ctypes.data, flag.ctypes.data, threads)
    ax = p.elec_axis
    L = float(p.box[ax])
    gcut = float(p.contact_gap)
    pts = [[] for _ in range(n)]  # (t, kind, index)
    for k in range(nc):
        if flag[k] != 0:
            continue
        pts[int(I[k])].append((float(tI[k]), 0, k))
        pts[int(J[k])].append((float(tJ[k]), 1, k))
    esrc, esnk = np.zeros(n, bool), np.zeros(n, bool)
    egs, egk = np.zeros(n), np.zeros(n)
    for i in range(n):
        hl = float(r[i, 6]) * 0.5
        ca = r[i, 3 + ax]
        ext = abs(hl * float(ca))
        pos, rad = float(r[i, ax]), float(r[i, 7])
        gs = (pos - ext) - rad
        gk = (L - (pos + ext)) - rad
        egs[i], egk[i] = gs, gk
        if gs <= gcut:
            esrc[i] = True
            pts[i].append((-hl if ca >= 0 else hl, 2, i))
        if gk <= gcut:
            esnk[i] = True
            pts[i].append((hl if ca >= 0 else -hl, 3, i))
    node_rod, node_t, sa, sb, sL = [], [], [], [], []
    edge_a = np.zeros(nc, np.uint32)
    edge_b = np.zeros(nc, np.uint32)
    src_node = np.zeros(n, np.uint32)
    snk_node = np.zeros(n, np.uint32)
    for i in range(n):
        ps = sorted(pts[i])
        if not ps:
            continue
        rad = float(r[i, 7])
        groups = [[ps[0]]]
        for q in range(1, len(ps)):
            if ps[q][0] - ps[q - 1][0] <= rad:
                groups[-1].append(ps[q])
            else:
                groups.append([ps[q]])
        prev = None
        for g in groups:
            tot = 0.0
            for x in g:
                tot += x[0]
            tm = tot / float(len(g))
            node = len(node_rod)
            node_rod.append(i)
            node_t.append(tm)
            if prev is not None:
                sa.append(node - 1)
                sb.append(node)
                sL.append((tm - prev) / rad)
            prev = tm
            for (_, kind, idx) in g:
                if kind == 0:
                    edge_a[idx] = node
                elif kind == 1:
                    edge_b[idx] = node
                elif kind == 2:
                    src_node[idx] = node
                else:
                    snk_node[idx] = node
    kept = np.flatnonzero(flag[:nc] == 0)
    elec = [(int(src_node[i]), egs[i], 0) for i in range(n) if esrc[i]] + \
           [(int(snk_node[i]), egk[i], 1) for i in range(n) if esnk[i]]
    return {
        "flag": flag[:nc].copy(), "tI": tI[:nc].copy(), "tJ": tJ[:nc].copy(),
        "node_rod": np.array(node_rod, np.uint32), "node_t": np.array(node_t),
        "seg_a": np.array(sa, np.uint32), "seg_b": np.array(sb, np.uint32), "seg_L": np.array(sL),
        "edge_a": edge_a[kept], "edge_b": edge_b[kept], "edge_contact": kept.astype(np.uint32),
        "elec_node": np.array([e[0] for e in elec], np.uint32),
        "elec_gap": np.array([e[1] for e in elec]), "elec_side": np.array([e[2] for e in elec], np.int32),
        "n_points": sum(len(x) for x in pts),
    }
