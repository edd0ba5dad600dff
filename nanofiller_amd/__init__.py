"""Python binding of the MI355X NanoFiller library (libnanofiller.so).

A thin ctypes layer over the C ABI in include/nanofiller.h, mirroring the
reference's host interface: the simPrms parameter state and keyword file
(inc/simPrms.h:57-178, src/simPrms.cpp:212-327), generation (cudaMain,
src/cudaMain.cu:61), contact arrays (src/get_contact_arrays.cu:12-140),
statistics (src/statistics_collection.cu), network extraction
(src/conductivity.cu:140) and the VTK / COO writers.

The compute path is the HIP library; there is no Python or CPU fallback.
Loading fails loudly if libnanofiller.so has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnanofiller.so")

NF_OK = 0
NF_E_INVALID, NF_E_NOMEM, NF_E_HIP, NF_E_IO = -1, -2, -3, -4
NF_E_OVERFLOW, NF_E_STATE, NF_E_NODEVICE, NF_E_CAPACITY = -5, -6, -7, -8
NF_LEN_CONST, NF_LEN_LOGNORMAL = 0, 1
NF_ORT_ISO, NF_ORT_VMF = 0, 1
NF_DEN_UNIFORM, NF_DEN_AGGLOM, NF_DEN_MESH = 0, 1, 2
NF_PREC_FP64, NF_PREC_FP32_RECHECK = 0, 1
NF_STATUS_RUNNING, NF_STATUS_TARGET, NF_STATUS_MAX_ROUNDS, NF_STATUS_STALLED, NF_STATUS_OVERFLOW = 0, 1, 2, 3, 4
STATUS = {0: "running", 1: "target", 2: "max_rounds", 3: "stalled", 4: "overflow"}
NF_OPT_GRAPH, NF_OPT_PROFILE, NF_OPT_DIAG, NF_OPT_CONTACT_PREC = 1, 2, 3, 4
NF_OPT_CONTACT_ENGINE, NF_OPT_FOOTPRINT_FINE, NF_OPT_FOOTPRINT_OCC, NF_OPT_POST_INSTANCE = 5, 6, 7, 8
NF_OPT_SLOT_SCALE = 9
NF_OPT_WAVES_OLD, NF_OPT_WAVES_NEW, NF_OPT_PART_SHIFT, NF_OPT_SPLIT_MIN = 10, 11, 12, 13
NF_OPT_TEST_ALL_PAIRS = 14
NF_ENGINE_FOOTPRINT, NF_ENGINE_CELLS = 0, 1
KERNEL_CLASSES = ["gen", "bin", "reject_old", "reject_new", "commit", "control", "contact", "other"]


class NFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nanofiller error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """nf_params (include/nanofiller.h)."""

    _fields_ = [
        ("box", C.c_double * 3),
        ("n_rods", C.c_uint32),
        ("seed", C.c_uint32),
        ("stream", C.c_uint32),
        ("len_law", C.c_int32),
        ("len", C.c_double),
        ("len_sigma", C.c_double),
        ("len_min", C.c_double),
        ("len_max", C.c_double),
        ("radius", C.c_double),
        ("sep", C.c_double),
        ("soft_frac", C.c_double),
        ("ort_law", C.c_int32),
        ("kappa", C.c_double),
        ("axis", C.c_double * 3),
        ("den_law", C.c_int32),
        ("n_agg", C.c_int32),
        ("agg_sigma", C.c_double),
        ("agg_peak", C.c_double),
        ("vol_frac", C.c_double),
        ("contact_gap", C.c_double),
        ("max_rounds", C.c_uint32),
        ("stall_rounds", C.c_uint32),
        ("min_batch", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("cell_size", C.c_double),
        ("precision", C.c_int32),
        ("elec_axis", C.c_int32),
        ("tunnel_lambda", C.c_double),
        ("tunnel_g0", C.c_double),
        ("write_state", C.c_int32),
        ("den_ext", C.c_int32 * 3),
        ("den_mesh", C.POINTER(C.c_float)),
    ]

    def copy(self) -> "Params":
        p = Params()
        C.pointer(p)[0] = self
        if hasattr(self, "_mesh"):
            p._mesh = self._mesh  # keeps the density map the pointer refers to alive
        return p

    def set_density_mesh(self, values: np.ndarray) -> None:
        """Density map law: values[ix, iy, iz] in [0, 1] over the box."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        assert v.ndim == 3
        self._mesh = v
        self.den_law = NF_DEN_MESH
        for i in range(3):
            self.den_ext[i] = v.shape[i]
        self.den_mesh = v.ctypes.data_as(C.POINTER(C.c_float))

    def to_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            if name == "den_mesh":
                continue
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class GenStats(C.Structure):
    """nf_gen_stats (include/nanofiller.h)."""

    _fields_ = [
        ("n_target", C.c_uint32),
        ("n_accepted", C.c_uint32),
        ("rounds", C.c_uint32),
        ("status", C.c_int32),
        ("candidates", C.c_uint64),
        ("proposals", C.c_uint64),
        ("pair_tests", C.c_uint64),
        ("rechecks", C.c_uint64),
        ("rejected_old", C.c_uint64),
        ("rejected_new", C.c_uint64),
        ("wall_ms", C.c_double),
        ("kern_ms", C.c_double * 8),
        ("max_abs_d32_d64", C.c_double),
        ("tile_bytes", C.c_uint64),
        ("tile_bytes_new", C.c_uint64),
        ("rounds_launched", C.c_uint64),
        ("slot_growths", C.c_uint32),
        ("reserved", C.c_uint32),
    ]

    def to_dict(self) -> dict:
        d = {n: getattr(self, n) for n, _ in self._fields_ if n != "kern_ms"}
        d["kern_ms"] = {k: self.kern_ms[i] for i, k in enumerate(KERNEL_CLASSES)}
        d["status"] = STATUS.get(self.status, str(self.status))
        return d


class HistSpec(C.Structure):
    _fields_ = [
        ("bins", C.c_int32),
        ("centre_axis", C.c_int32),
        ("pol_axis", C.c_double * 3),
        ("phi_bins", C.c_int32),
        ("theta_bins", C.c_int32),
    ]


class NetworkInfo(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32),
        ("n_components", C.c_uint32),
        ("largest", C.c_uint32),
        ("percolating", C.c_uint32),
        ("nnz", C.c_uint64),
        ("n_edges", C.c_uint64),
        ("n_source", C.c_uint32),
        ("n_sink", C.c_uint32),
    ]

    def to_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class StateHeader(C.Structure):
    """nf_state_header (include/nanofiller.h)."""

    _fields_ = [
        ("box", C.c_float * 3),
        ("edge", C.c_float * 3),
        ("ext", C.c_int32 * 3),
        ("len", C.c_float),
        ("radius", C.c_float),
        ("slots", C.c_uint32),
        ("n_rods", C.c_uint32),
        ("tol", C.c_float),
    ]

    def to_dict(self) -> dict:
        return {"box": tuple(self.box), "edge": tuple(self.edge), "ext": tuple(self.ext),
                "len": self.len, "radius": self.radius, "slots": self.slots,
                "n_rods": self.n_rods, "tol": self.tol}


class NodeInfo(C.Structure):
    """nf_node_info (include/nanofiller.h)."""

    _fields_ = [
        ("n_points", C.c_uint32),
        ("n_nodes", C.c_uint32),
        ("n_segments", C.c_uint32),
        ("n_edges", C.c_uint32),
        ("n_screened", C.c_uint32),
        ("n_source", C.c_uint32),
        ("n_sink", C.c_uint32),
        ("pad", C.c_uint32),
        ("device_ms", C.c_double),
    ]

    def to_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad"}


class ReduceInfo(C.Structure):
    """nf_reduce_info (include/nanofiller.h)."""

    _fields_ = [
        ("n_before", C.c_uint32),
        ("n_after", C.c_uint32),
        ("rounds", C.c_uint32),
        ("pad", C.c_uint32),
        ("device_ms", C.c_double),
    ]

    def to_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad"}


class ConductanceInfo(C.Structure):
    """nf_conductance_info (include/nanofiller.h)."""

    _fields_ = [
        ("conductance", C.c_double),
        ("current_sink", C.c_double),
        ("residual", C.c_double),
        ("solve_ms", C.c_double),
        ("unknowns", C.c_uint32),
        ("iterations", C.c_uint32),
        ("converged", C.c_int32),
        ("pad", C.c_uint32),
    ]

    def to_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad"}


_lib_handle = None

# (name, restype, argtypes) of every exported entry point
_SIGS = [
    ("nf_params_default", None, [C.POINTER(Params)]),
    ("nf_params_load", C.c_int, [C.c_char_p, C.POINTER(Params)]),
    ("nf_params_normalize", C.c_int, [C.POINTER(Params)]),
    ("nf_agglomerate_centres", C.c_int, [C.POINTER(Params), C.c_void_p, C.c_int32]),
    ("nf_create", C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    ("nf_create_batch", C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(C.c_void_p)]),
    ("nf_destroy", None, [C.c_void_p]),
    ("nf_last_error", C.c_char_p, [C.c_void_p]),
    ("nf_num_instances", C.c_int, [C.c_void_p]),
    ("nf_generate", C.c_int, [C.c_void_p, C.POINTER(GenStats)]),
    ("nf_get_rods", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
    ("nf_set_rods", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint32]),
    ("nf_get_params", C.c_int, [C.c_void_p, C.c_int, C.POINTER(Params)]),
    ("nf_expand", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    ("nf_set_milestones", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_char_p]),
    ("nf_params_milestones", C.c_int, [C.c_char_p, C.c_void_p, C.c_int32]),
    ("nf_contacts", C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_uint64)]),
    ("nf_get_contacts", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nf_get_contact_gap64", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nf_get_contact_area", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nf_contacts_engine", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    ("nf_contacts_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_double), C.c_void_p, C.POINTER(C.c_uint64)]),
    ("nf_centre_distance", C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_void_p,
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]),
    ("nf_vdw", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double),
                         C.POINTER(C.c_uint64)]),
    ("nf_histograms", C.c_int, [C.c_void_p, C.POINTER(HistSpec), C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p]),
    ("nf_histograms_noalloc", C.c_int, [C.c_void_p, C.POINTER(HistSpec), C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p]),
    ("nf_network", C.c_int, [C.c_void_p, C.POINTER(NetworkInfo)]),
    ("nf_get_network", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nf_write_vtk", C.c_int, [C.c_void_p, C.c_char_p]),
    ("nf_write_coo", C.c_int, [C.c_void_p, C.c_char_p]),
    ("nf_reduce", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ReduceInfo)]),
    ("nf_get_reduce_kept", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nf_conductance", C.c_int, [C.c_void_p, C.c_double, C.c_uint32, C.POINTER(ConductanceInfo),
                                 C.c_void_p]),
    ("nf_read_density_mesh", C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nf_write_density_mesh", C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("nf_node_network", C.c_int, [C.c_void_p, C.POINTER(NodeInfo)]),
    ("nf_get_nodes", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nf_get_segments", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_uint64]),
    ("nf_get_node_edges", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                    C.POINTER(C.c_uint64)]),
    ("nf_get_node_electrodes", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                         C.POINTER(C.c_uint64)]),
    ("nf_write_node_coo", C.c_int, [C.c_void_p, C.c_char_p]),
    ("nf_save_state", C.c_int, [C.c_void_p, C.c_char_p]),
    ("nf_read_state_header", C.c_int, [C.c_char_p, C.POINTER(StateHeader)]),
    ("nf_load_state", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_uint32)]),
    ("nf_version", C.c_char_p, []),
    ("nf_set_option", C.c_int, [C.c_void_p, C.c_int, C.c_int64]),
    ("nf_selftest", C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("nf_selftest_pairs", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_double,
                                    C.c_double, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
]


def lib() -> C.CDLL:
    """The loaded HIP library; raises if it has not been built."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m nanofiller_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib_handle = L
    return _lib_handle


def params_milestones(path: str) -> list[float]:
    """The VOLUME_PERCENT_LIST of a parameter file (percent)."""
    n = lib().nf_params_milestones(path.encode(), None, 0)
    if n < 0:
        _check(n)
    out = np.zeros(max(n, 1), np.float64)
    lib().nf_params_milestones(path.encode(), out.ctypes.data, n)
    return [float(x) for x in out[:n]]


def exported_symbols() -> list[str]:
    return [s[0] for s in _SIGS]


def _check(rc: int, h=None) -> None:
    if rc != NF_OK:
        msg = lib().nf_last_error(h)
        raise NFError(rc, msg.decode() if msg else "")


def default_params(**kw) -> Params:
    """simPrms defaults (src/simPrms.cpp:22-117) with field overrides."""
    p = Params()
    lib().nf_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "den_mesh":
            p.set_density_mesh(v)
        elif k in ("box", "axis"):
            arr = getattr(p, k)
            for i in range(3):
                arr[i] = float(v[i])
        else:
            setattr(p, k, v)
    return p


def read_density_mesh(path: str):
    """Density map file (read_density format): (dims, values[ix, iy, iz])."""
    dim = (C.c_double * 3)()
    ext = (C.c_int32 * 3)()
    _check(lib().nf_read_density_mesh(path.encode(), dim, ext, None, 0))
    v = np.zeros(tuple(ext), np.float32)
    _check(lib().nf_read_density_mesh(path.encode(), dim, ext, v.ctypes.data, v.size))
    return tuple(dim), v


def write_density_mesh(path: str, dims, values: np.ndarray) -> None:
    v = np.ascontiguousarray(values, dtype=np.float32)
    dim = (C.c_double * 3)(*[float(x) for x in dims])
    ext = (C.c_int32 * 3)(*v.shape)
    _check(lib().nf_write_density_mesh(path.encode(), dim, ext, v.ctypes.data))


def agglomerate_centres(p: Params) -> np.ndarray:
    """(n_agg, 3) centres of the agglomerate density law of p."""
    n = lib().nf_agglomerate_centres(C.byref(p), None, 0)
    if n < 0:
        _check(n)
    out = np.zeros((max(n, 0), 3), np.float64)
    if n > 0:
        rc = lib().nf_agglomerate_centres(C.byref(p), out.ctypes.data, n)
        if rc < 0:
            _check(rc)
    return out


def load_params(path: str, base: Params | None = None) -> Params:
    """Keyword parameter file (simPrms::set_ParamsFromFile)."""
    p = base.copy() if base is not None else default_params()
    _check(lib().nf_params_load(path.encode(), C.byref(p)))
    return p


def normalize(p: Params) -> Params:
    q = p.copy()
    _check(lib().nf_params_normalize(C.byref(q)))
    return q


@dataclass
class Contacts:
    row_ptr: np.ndarray  # uint32 [N+1]
    col: np.ndarray      # uint32 [C]
    gap: np.ndarray      # float32 [C]
    image: np.ndarray    # int32 [C] (3 signed bytes)

    @property
    def I(self) -> np.ndarray:  # noqa: E743
        return np.repeat(np.arange(len(self.row_ptr) - 1, dtype=np.uint32), np.diff(self.row_ptr))

    @property
    def n(self) -> int:
        return int(self.row_ptr[-1])


class Packing:
    """One handle: a single packing or a batch of independent packings."""

    def __init__(self, params, options: dict | None = None, _handle=None):
        if _handle is not None:
            self._h = _handle  # a handle the library made (nf_expand)
            self.n_inst = 1
            self.post_inst = 0
            return
        plist = [params] if isinstance(params, Params) else list(params)
        arr = (Params * len(plist))(*plist)
        h = C.c_void_p()
        _check(lib().nf_create_batch(arr, len(plist), C.byref(h)))
        self._h = h
        self.n_inst = len(plist)
        self.post_inst = 0
        for k, v in (options or {}).items():
            self.set_option(k, v)

    def params(self, inst: int = 0) -> Params:
        """The instance's normalised parameters."""
        p = Params()
        _check(lib().nf_get_params(self._h, inst, C.byref(p)), self._h)
        return p

    def expand(self, direction: int) -> "Packing":
        """A new packing: this one with periodic copies along the axis of
        `direction` (bit mask 1 = x, 2 = y, 4 = z), nf_expand."""
        g = C.c_void_p()
        _check(lib().nf_expand(self._h, int(direction), C.byref(g)), self._h)
        return Packing(None, _handle=g)

    def set_milestones(self, vf_percent, directory: str = "") -> None:
        """State dumps at volume fractions (percent) during generate()."""
        v = np.ascontiguousarray(vf_percent, np.float64)
        _check(lib().nf_set_milestones(self._h, v.ctypes.data if len(v) else None, len(v),
                                       directory.encode()), self._h)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().nf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_option(self, key: int, value: int) -> None:
        _check(lib().nf_set_option(self._h, int(key), int(value)), self._h)
        if key == NF_OPT_POST_INSTANCE:
            self.post_inst = int(value)

    def select(self, inst: int) -> "Packing":
        """Post-process instance `inst` of a batch (NF_OPT_POST_INSTANCE)."""
        self.set_option(NF_OPT_POST_INSTANCE, inst)
        return self

    def generate(self) -> list[GenStats]:
        st = (GenStats * self.n_inst)()
        _check(lib().nf_generate(self._h, st), self._h)
        return list(st)

    def rods(self, inst: int = 0) -> np.ndarray:
        n = C.c_uint32()
        _check(lib().nf_get_rods(self._h, inst, None, 0, C.byref(n)), self._h)
        out = np.zeros((n.value, 8), dtype=np.float32)
        _check(lib().nf_get_rods(self._h, inst, out.ctypes.data, n.value, C.byref(n)), self._h)
        return out

    def set_rods(self, rods: np.ndarray, inst: int = 0) -> None:
        r = np.ascontiguousarray(rods, dtype=np.float32)
        assert r.ndim == 2 and r.shape[1] == 8
        _check(lib().nf_set_rods(self._h, inst, r.ctypes.data, r.shape[0]), self._h)

    def contacts(self, gap: float) -> Contacts:
        nc = C.c_uint64()
        _check(lib().nf_contacts(self._h, float(gap), C.byref(nc)), self._h)
        nrod = C.c_uint32()
        _check(lib().nf_get_rods(self._h, self.post_inst, None, 0, C.byref(nrod)), self._h)
        row = np.zeros(nrod.value + 1, dtype=np.uint32)
        col = np.zeros(nc.value, dtype=np.uint32)
        g = np.zeros(nc.value, dtype=np.float32)
        im = np.zeros(nc.value, dtype=np.int32)
        _check(lib().nf_get_contacts(self._h, row.ctypes.data, col.ctypes.data, g.ctypes.data,
                                     im.ctypes.data, nc.value), self._h)
        return Contacts(row, col, g, im)

    def contact_gap64(self, n: int) -> np.ndarray:
        out = np.zeros(n, dtype=np.float64)
        _check(lib().nf_get_contact_gap64(self._h, out.ctypes.data, n), self._h)
        return out

    def contact_area(self, n: int) -> np.ndarray:
        """Contact areas S of the last contacts() list, in its order."""
        out = np.zeros(n, dtype=np.float64)
        _check(lib().nf_get_contact_area(self._h, out.ctypes.data, n), self._h)
        return out

    def contact_stats(self) -> dict:
        t, r, m, tb = C.c_uint64(), C.c_uint64(), C.c_double(), C.c_uint64()
        ms = (C.c_float * 4)()
        _check(lib().nf_contacts_stats(self._h, C.byref(t), C.byref(r), C.byref(m), ms, C.byref(tb)),
               self._h)
        return {"pair_tests": t.value, "rechecks": r.value, "max_abs_d32_d64": m.value,
                "tile_bytes": tb.value,
                "kernel_ms": {"bin": ms[0], "pairs": ms[1], "csr": ms[2], "sort": ms[3]}}

    def contact_engine(self) -> dict:
        """Engine that found the last contact list (after any fallback), and
        the footprint engine's home cell edge, fine cells per axis and work
        units redone one cell at a time."""
        e, edge, f, sp = C.c_int(), C.c_double(), C.c_int(), C.c_uint64()
        _check(lib().nf_contacts_engine(self._h, C.byref(e), C.byref(edge), C.byref(f), C.byref(sp)),
               self._h)
        return {"engine": e.value, "home_edge": edge.value, "fine": f.value, "split_units": sp.value}

    def histograms(self, bins=64, centre_axis=0, pol_axis=(1.0, 0.0, 0.0), mutual=True,
                   phi_bins=0, theta_bins=0, noalloc: bool = False) -> dict:
        """Exact int64 histograms; noalloc=True calls nf_histograms_noalloc
        (scratch in the library's static device array, no allocation)."""
        spec = HistSpec()
        spec.bins = bins
        spec.centre_axis = centre_axis
        for i in range(3):
            spec.pol_axis[i] = pol_axis[i]
        spec.phi_bins = phi_bins
        spec.theta_bins = theta_bins
        c = np.zeros(bins, np.int64)
        p = np.zeros(bins, np.int64)
        m = np.zeros(bins, np.int64) if mutual else None
        pt = np.zeros(max(phi_bins * theta_bins, 1), np.int64) if phi_bins else None
        fn = lib().nf_histograms_noalloc if noalloc else lib().nf_histograms
        _check(fn(self._h, C.byref(spec), c.ctypes.data, p.ctypes.data,
                  m.ctypes.data if m is not None else None,
                  pt.ctypes.data if pt is not None else None), self._h)
        out = {"centre": c, "polar": p}
        if m is not None:
            out["mutual"] = m
        if pt is not None:
            out["phitheta"] = pt.reshape(theta_bins, phi_bins)
        return out

    def centre_distance(self, bins: int, width: float) -> dict:
        """Centre-centre distance histogram (statKernel_CENT) with exact
        fixed-point sums; mean and variance derived from them."""
        counts = np.zeros(bins, np.int64)
        n, s1, s2 = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().nf_centre_distance(self._h, bins, width, counts.ctypes.data, C.byref(n),
                                        C.byref(s1), C.byref(s2)), self._h)
        npairs = n.value
        mean = s1.value / 16777216.0 / npairs if npairs else 0.0
        var = s2.value / 65536.0 / npairs - mean * mean if npairs else 0.0
        return {"counts": counts, "pairs": npairs, "sum_d_q24": s1.value, "sum_d2_q16": s2.value,
                "mean": mean, "var": var}

    def vdw(self, A: float = 15.2, B: float = 24100.0) -> tuple[float, int]:
        """Van der Waals energy over the last contact list (statKernel_VdW_naive)."""
        e, n = C.c_double(), C.c_uint64()
        _check(lib().nf_vdw(self._h, A, B, C.byref(e), C.byref(n)), self._h)
        return e.value, n.value

    def network(self) -> tuple[dict, dict]:
        info = NetworkInfo()
        _check(lib().nf_network(self._h, C.byref(info)), self._h)
        n = info.n_nodes - 2
        labels = np.zeros(n, np.uint32)
        row = np.zeros(info.n_nodes + 1, np.uint32)
        col = np.zeros(info.nnz, np.uint32)
        val = np.zeros(info.nnz, np.float64)
        _check(lib().nf_get_network(self._h, labels.ctypes.data, row.ctypes.data, col.ctypes.data,
                                    val.ctypes.data, info.nnz), self._h)
        return info.to_dict(), {"labels": labels, "row_ptr": row, "col": col, "val": val}

    def node_network(self) -> tuple[dict, dict]:
        """Node-level segment network of the last contacts(): (info, arrays)."""
        info = NodeInfo()
        _check(lib().nf_node_network(self._h, C.byref(info)), self._h)
        nn, ns = info.n_nodes, info.n_segments
        rod = np.zeros(nn, np.uint32)
        t = np.zeros(nn, np.float64)
        _check(lib().nf_get_nodes(self._h, rod.ctypes.data, t.ctypes.data, nn), self._h)
        sa = np.zeros(ns, np.uint32)
        sb = np.zeros(ns, np.uint32)
        sl = np.zeros(ns, np.float64)
        sarea = np.zeros(ns, np.float64)
        _check(lib().nf_get_segments(self._h, sa.ctypes.data, sb.ctypes.data, sl.ctypes.data,
                                     sarea.ctypes.data, ns), self._h)
        cnt = C.c_uint64()
        _check(lib().nf_get_node_edges(self._h, None, None, None, 0, C.byref(cnt)), self._h)
        ea = np.zeros(cnt.value, np.uint32)
        eb = np.zeros(cnt.value, np.uint32)
        ek = np.zeros(cnt.value, np.uint32)
        _check(lib().nf_get_node_edges(self._h, ea.ctypes.data, eb.ctypes.data, ek.ctypes.data,
                                       cnt.value, C.byref(cnt)), self._h)
        _check(lib().nf_get_node_electrodes(self._h, None, None, None, 0, C.byref(cnt)), self._h)
        en = np.zeros(cnt.value, np.uint32)
        eg = np.zeros(cnt.value, np.float64)
        es = np.zeros(cnt.value, np.int32)
        _check(lib().nf_get_node_electrodes(self._h, en.ctypes.data, eg.ctypes.data, es.ctypes.data,
                                            cnt.value, C.byref(cnt)), self._h)
        return info.to_dict(), {"node_rod": rod, "node_t": t, "seg_a": sa, "seg_b": sb, "seg_L": sl,
                                "seg_A": sarea, "edge_a": ea, "edge_b": eb, "edge_contact": ek,
                                "elec_node": en, "elec_gap": eg, "elec_side": es}

    def write_node_coo(self, prefix: str = "") -> None:
        _check(lib().nf_write_node_coo(self._h, prefix.encode()), self._h)

    def reduce(self, min_cont: int = 2) -> tuple[dict, np.ndarray]:
        """Dead-end pruning of the last contacts(): (info, old id of every
        kept rod).  The kept rods become the packing; call contacts() again."""
        info = ReduceInfo()
        _check(lib().nf_reduce(self._h, int(min_cont), C.byref(info)), self._h)
        ids = np.zeros(info.n_after, np.uint32)
        _check(lib().nf_get_reduce_kept(self._h, ids.ctypes.data, info.n_after), self._h)
        return info.to_dict(), ids

    def conductance(self, tol: float = 1e-10, max_iter: int = 20000, potentials: bool = False):
        """DC conductance between the electrodes of the last network():
        (info dict, rod potentials or None)."""
        info = ConductanceInfo()
        n = C.c_uint32()
        _check(lib().nf_get_rods(self._h, self.post_inst, None, 0, C.byref(n)), self._h)
        v = np.zeros(max(n.value, 1), np.float64) if potentials else None
        _check(lib().nf_conductance(self._h, tol, max_iter, C.byref(info),
                                    v.ctypes.data if v is not None else None), self._h)
        return info.to_dict(), (v[: n.value] if v is not None else None)

    def write_vtk(self, path: str) -> None:
        _check(lib().nf_write_vtk(self._h, path.encode()), self._h)

    def write_coo(self, prefix: str = "") -> None:
        _check(lib().nf_write_coo(self._h, prefix.encode()), self._h)

    def save_state(self, path: str) -> None:
        """Instance 0 as the reference's binary state file (output.bin)."""
        _check(lib().nf_save_state(self._h, path.encode()), self._h)

    def load_state(self, path: str) -> int:
        """Replaces the post-processing instance's rods with a state file's; returns the count."""
        n = C.c_uint32()
        _check(lib().nf_load_state(self._h, path.encode(), C.byref(n)), self._h)
        return n.value


def read_state_header(path: str) -> dict:
    """Header of a binary state file (checked against the file size)."""
    hd = StateHeader()
    _check(lib().nf_read_state_header(path.encode(), C.byref(hd)), None)
    return hd.to_dict()


SELFTEST = {"sqrt": 0, "exp": 1, "log": 2, "div": 3, "u01": 4, "hw_sqrt": 5, "copy": 6}


def selftest(which: str, x: np.ndarray) -> np.ndarray:
    """Evaluate one FP64 device building block element-wise on the GPU."""
    k = SELFTEST[which]
    a = np.ascontiguousarray(x, dtype=np.float64)
    n = a.size // 2 if k == 3 else a.size
    out = np.zeros(n, np.float64)
    _check(lib().nf_selftest(k, a.ctypes.data, out.ctypes.data, n))
    return out


def selftest_pairs(ra: np.ndarray, rb: np.ndarray, shift: np.ndarray, contact: bool, sep_or_gap: float,
                   ew: float, soft_frac: float = 1.0) -> dict:
    """The FP32 screen and band of the pair passes evaluated on the GPU for
    each pair (rods as rows of 8 floats, rod b moved by its shift)."""
    a = np.ascontiguousarray(ra, np.float32)
    b = np.ascontiguousarray(rb, np.float32)
    sh = np.ascontiguousarray(shift, np.float64)
    n = a.shape[0]
    scr = np.zeros(n, np.int32)
    dec = np.zeros(n, np.int32)
    wi = np.zeros(n, np.int32)
    _check(lib().nf_selftest_pairs(a.ctypes.data, b.ctypes.data, sh.ctypes.data, n, int(contact),
                                   float(sep_or_gap), float(soft_frac), float(ew), scr.ctypes.data,
                                   dec.ctypes.data, wi.ctypes.data))
    return {"screen": scr, "decided": dec, "within": wi.astype(bool)}


def version() -> str:
    return lib().nf_version().decode()
