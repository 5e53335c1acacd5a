"""Python binding of the MI355X shallow-water library (ctypes over the C ABI
declared in include/omega_swe.h).

The classes keep the reference's names and argument meaning
(mwarusz/omega_playground: PlanarHexagonalMesh, ShallowWaterParams,
ShallowWaterState, ShallowWaterModel, RK4Stepper, LSRKStepper and the
diagnostics integrals).  Host arrays use the reference's original entity
numbering and LayoutRight shapes; device buffers use the library's
space-filling-curve numbering.  Every call goes to libomega_swe.so; there is
no CPU fallback, and importing fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libomega_swe.so")

ORDER = {"original": 0, "hilbert": 1, "morton": 2}
ADD_REPLACE, ADD_INCREMENT = 0, 1
PV, KE, SSH, VDEL2, VDEL4, TDEL2, TDEL4, DRAG = 1, 2, 4, 8, 16, 32, 64, 128

_lib = None


class OmegaError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("f0", C.c_double), ("grav", C.c_double), ("drag_coeff", C.c_double),
        ("visc_del2", C.c_double), ("visc_del4", C.c_double), ("eddy_diff2", C.c_double),
        ("eddy_diff4", C.c_double), ("ntracers", C.c_int), ("disable_h_tendency", C.c_int),
        ("disable_vn_tendency", C.c_int),
    ]


class MpasMeshDesc(C.Structure):
    _ip = C.POINTER(C.c_int)
    _dp = C.POINTER(C.c_double)
    _fields_ = [
        ("ncells", C.c_int), ("nedges", C.c_int), ("nvertices", C.c_int), ("nlayers", C.c_int),
        ("max_edges", C.c_int), ("max_edges2", C.c_int),
        ("nedges_on_cell", _ip), ("edges_on_cell", _ip), ("vertices_on_cell", _ip),
        ("nedges_on_edge", _ip), ("edges_on_edge", _ip), ("cells_on_edge", _ip),
        ("vertices_on_edge", _ip), ("edges_on_vertex", _ip), ("cells_on_vertex", _ip),
        ("area_cell", _dp), ("dc_edge", _dp), ("dv_edge", _dp), ("weights_on_edge", _dp),
        ("area_triangle", _dp), ("kiteareas_on_vertex", _dp), ("mesh_scaling_del2", _dp),
        ("mesh_scaling_del4", _dp), ("max_level_edge_top", _ip),
        ("x_cell", _dp), ("y_cell", _dp), ("x_edge", _dp), ("y_edge", _dp), ("angle_edge", _dp),
        ("x_vertex", _dp), ("y_vertex", _dp),
    ]


ADDFN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                    C.c_double, C.c_void_p)


def lib():
    """Load libomega_swe.so (built by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OmegaError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                         "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp, i64p = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int64)
    pp = C.POINTER(C.c_void_p)
    sig = {
        "omega_last_error": ([], C.c_char_p),
        "omega_version": ([], C.c_char_p),
        "omega_mesh_create_planar_hex": ([C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, pp], C.c_int),
        "omega_mesh_create_mpas": ([C.POINTER(MpasMeshDesc), C.c_int, pp], C.c_int),
        "omega_mesh_destroy": ([vp], C.c_int),
        "omega_mesh_info": ([vp, ip, dp], C.c_int),
        "omega_mesh_host_array": ([vp, C.c_char_p, pp, ip, i64p, i64p], C.c_int),
        "omega_mesh_device_array": ([vp, C.c_char_p, pp, i64p], C.c_int),
        "omega_mesh_permutation": ([vp, C.c_int, ip], C.c_int),
        "omega_mesh_regular_hex": ([vp, ip], C.c_int),
        "omega_state_create": ([vp, C.c_int, pp], C.c_int),
        "omega_state_destroy": ([vp], C.c_int),
        "omega_state_upload": ([vp, dp, dp, dp, vp], C.c_int),
        "omega_state_download": ([vp, dp, dp, dp, vp], C.c_int),
        "omega_state_device_ptrs": ([vp, pp, pp, pp], C.c_int),
        "omega_state_copy": ([vp, vp, vp], C.c_int),
        "omega_model_create": ([vp, C.POINTER(Params), pp], C.c_int),
        "omega_model_create_linear": ([vp, C.POINTER(Params), C.c_double, pp], C.c_int),
        "omega_model_destroy": ([vp], C.c_int),
        "omega_model_set_f_vertex": ([vp, dp], C.c_int),
        "omega_model_set_additional_tendency": ([vp, ADDFN, vp], C.c_int),
        "omega_model_compute_tendency": ([vp, vp, vp, C.c_double, C.c_int, vp], C.c_int),
        "omega_model_set_terms": ([vp, C.c_int], C.c_int),
        "omega_model_compute_field_tendency": ([vp, C.c_int, vp, vp, C.c_int, vp], C.c_int),
        "omega_model_compute_aux": ([vp, vp, vp, vp, vp], C.c_int),
        "omega_model_compute_tendency_noalloc": ([vp, vp, vp, C.c_double, C.c_int, vp], C.c_int),
        "omega_model_compute_h_tendency": ([vp, vp, vp, vp, C.c_int, vp], C.c_int),
        "omega_model_compute_vn_tendency": ([vp, vp, vp, vp, C.c_int, vp], C.c_int),
        "omega_model_compute_tr_tendency": ([vp, vp, vp, vp, C.c_int, vp], C.c_int),
        "omega_model_aux_download": ([vp, C.c_char_p, dp, vp], C.c_int),
        "omega_stepper_coeffs": ([vp, ip, dp, dp, dp], C.c_int),
        "omega_model_get_terms": ([vp, ip], C.c_int),
        "omega_model_set_launch": ([vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "omega_model_get_launch": ([vp, ip, ip, ip, ip], C.c_int),
        "omega_model_autotune": ([vp, vp, C.c_int, dp, vp], C.c_int),
        "omega_model_set_variant": ([vp, C.c_int], C.c_int),
        "omega_model_kernel_variant": ([vp, ip], C.c_int),
        "omega_model_kernel_flags": ([vp, ip], C.c_int),
        "omega_model_set_kernel_forms": ([vp, C.c_int], C.c_int),
        "omega_stepper_create": ([vp, C.c_int, C.c_int, pp], C.c_int),
        "omega_stepper_destroy": ([vp], C.c_int),
        "omega_stepper_do_step": ([vp, C.c_double, C.c_double, vp, vp], C.c_int),
        "omega_stepper_run": ([vp, C.c_double, C.c_double, C.c_int, C.c_int, vp, vp], C.c_int),
        "omega_stepper_set_graph": ([vp, C.c_int], C.c_int),
        "omega_energy_integral": ([vp, vp, dp, vp], C.c_int),
        "omega_mass_integral": ([vp, vp, dp, vp], C.c_int),
        "omega_circulation_integral": ([vp, vp, dp, vp], C.c_int),
        "omega_add_host_field": ([vp, C.c_int, vp, dp, C.c_int, C.c_double, vp], C.c_int),
        "omega_synchronize": ([vp], C.c_int),
        "omega_timing_enable": ([C.c_int], C.c_int),
        "omega_timing_reset": ([], C.c_int),
        "omega_timing_read": ([C.c_char_p, dp, i64p], C.c_int),
        "omega_model_tendency_vjp": ([vp, vp, vp, vp, dp, vp], C.c_int),
        "omega_stepper_step_vjp": ([vp, C.c_double, C.c_double, vp, vp, dp, vp], C.c_int),
        "omega_tape_create": ([vp, C.c_int, C.c_int, C.c_int64, pp], C.c_int),
        "omega_tape_destroy": ([vp], C.c_int),
        "omega_tape_forward": ([vp, C.c_double, C.c_double, vp, vp], C.c_int),
        "omega_tape_reverse": ([vp, vp, dp, vp], C.c_int),
        "omega_tape_info": ([vp, ip, i64p, i64p], C.c_int),
        "omega_revolve_cost": ([C.c_int64, C.c_int], C.c_int64),
        "omega_misfit": ([vp, vp, vp, dp, vp, dp, vp], C.c_int),
        "omega_adjoint_run": ([vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int64, vp, vp,
                               dp, dp, vp, dp, vp], C.c_int),
        "omega_batch_create": ([C.c_int, C.c_int, pp], C.c_int),
        "omega_batch_destroy": ([vp], C.c_int),
        "omega_batch_add": ([vp, vp, vp, C.c_double, ip], C.c_int),
        "omega_batch_step": ([vp, vp], C.c_int),
        "omega_batch_run": ([vp, C.c_int, vp], C.c_int),
        "omega_batch_set_graph": ([vp, C.c_int], C.c_int),
        "omega_batch_set_launch": ([vp, C.c_int, C.c_int], C.c_int),
        "omega_batch_get_launch": ([vp, ip, ip], C.c_int),
        "omega_batch_autotune": ([vp, C.c_int, dp, vp], C.c_int),
        "omega_batch_set_persistent": ([vp, C.c_int], C.c_int),
        "omega_batch_get_persistent": ([vp, ip], C.c_int),
        "omega_batch_info": ([vp, ip, i64p, i64p], C.c_int),
        "omega_batch_adjoint_run": ([vp, C.c_double, C.c_int, C.c_int, C.c_int64, pp, dp, dp, pp,
                                     dp, vp], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        raise OmegaError(f"omega error {rc}: {lib().omega_last_error().decode()}")


def _dp(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "need C-contiguous float64"
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _stream(s) -> C.c_void_p:
    if s is None:
        return C.c_void_p(0)
    if hasattr(s, "cuda_stream"):  # torch.cuda.Stream
        return C.c_void_p(s.cuda_stream)
    return C.c_void_p(int(s))


# ---------------------------------------------------------------------------
class MPASMesh:
    """Device mesh (MPASMesh, src/mesh/mpas_mesh.hpp:7-71).  Host fields are
    available as attributes in the original numbering (mesh.x_cell, ...)."""

    _h = None

    def _init_info(self):
        counts = (C.c_int * 6)()
        reals = (C.c_double * 3)()
        check(lib().omega_mesh_info(self._h, counts, reals))
        (self.ncells, self.nedges, self.nvertices, self.nlayers, self.nx, self.ny) = list(counts)
        self.dc, self.period_x, self.period_y = list(reals)
        self._perm = {}

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        ptr, dt, n0, n1 = C.c_void_p(), C.c_int(), C.c_int64(), C.c_int64()
        rc = lib().omega_mesh_host_array(self._h, name.encode(), C.byref(ptr), C.byref(dt),
                                         C.byref(n0), C.byref(n1))
        if rc != 0:
            raise AttributeError(name)
        ct = C.c_int if dt.value == 0 else C.c_double
        buf = (ct * (n0.value * n1.value)).from_address(ptr.value)
        a = np.frombuffer(buf, dtype=np.int32 if dt.value == 0 else np.float64)
        return a.reshape(n0.value, n1.value) if n1.value != 1 else a

    def permutation(self, kind: int) -> np.ndarray:
        """new2old map: 0 cells, 1 edges, 2 vertices."""
        if kind not in self._perm:
            n = (self.ncells, self.nedges, self.nvertices)[kind]
            out = np.empty(n, dtype=np.int32)
            check(lib().omega_mesh_permutation(self._h, kind, out.ctypes.data_as(C.POINTER(C.c_int))))
            self._perm[kind] = out
        return self._perm[kind]

    @property
    def regular_hex(self) -> bool:
        """True when the pattern-specialised kernels can run on this mesh."""
        v = C.c_int()
        check(lib().omega_mesh_regular_hex(self._h, C.byref(v)))
        return bool(v.value)

    def device_array(self, name: str):
        p, n = C.c_void_p(), C.c_int64()
        check(lib().omega_mesh_device_array(self._h, name.encode(), C.byref(p), C.byref(n)))
        return p.value, n.value

    def add_host_field(self, kind: int, dst_ptr: int, src: np.ndarray, alpha=1.0, stream=None):
        src = np.ascontiguousarray(src, dtype=np.float64)
        ncols = 1 if src.ndim == 1 else src.shape[1]
        check(lib().omega_add_host_field(self._h, kind, C.c_void_p(dst_ptr), _dp(src), ncols,
                                         float(alpha), _stream(stream)))

    def __del__(self):
        if self._h is not None and _lib is not None:
            _lib.omega_mesh_destroy(self._h)
            self._h = None

    @classmethod
    def from_arrays(cls, arrays: dict, nlayers: int, ordering: str = "hilbert"):
        """Build from MPAS arrays (reference names, original numbering)."""
        self = cls.__new__(cls)
        keep = {}

        def ia(name):
            a = np.ascontiguousarray(arrays[name], dtype=np.int32)
            keep[name] = a
            return a.ctypes.data_as(C.POINTER(C.c_int))

        def da(name, required=True):
            if name not in arrays or arrays[name] is None:
                if required:
                    raise KeyError(name)
                return None
            a = np.ascontiguousarray(arrays[name], dtype=np.float64)
            keep[name] = a
            return a.ctypes.data_as(C.POINTER(C.c_double))

        eoc = np.asarray(arrays["edges_on_cell"])
        eoe = np.asarray(arrays["edges_on_edge"])
        d = MpasMeshDesc()
        d.ncells, d.nedges = eoc.shape[0], eoe.shape[0]
        d.nvertices = np.asarray(arrays["edges_on_vertex"]).shape[0]
        d.nlayers = nlayers
        d.max_edges, d.max_edges2 = eoc.shape[1], eoe.shape[1]
        for f in ("nedges_on_cell", "edges_on_cell", "vertices_on_cell", "nedges_on_edge",
                  "edges_on_edge", "cells_on_edge", "vertices_on_edge", "edges_on_vertex",
                  "cells_on_vertex"):
            setattr(d, f, ia(f))
        for f in ("area_cell", "dc_edge", "dv_edge", "weights_on_edge", "area_triangle",
                  "kiteareas_on_vertex"):
            setattr(d, f, da(f))
        for f in ("mesh_scaling_del2", "mesh_scaling_del4", "x_cell", "y_cell", "x_edge", "y_edge",
                  "angle_edge", "x_vertex", "y_vertex"):
            setattr(d, f, da(f, required=False))
        d.max_level_edge_top = ia("max_level_edge_top") if "max_level_edge_top" in arrays else None
        h = C.c_void_p()
        check(lib().omega_mesh_create_mpas(C.byref(d), ORDER[ordering], C.byref(h)))
        self._h = h
        self._init_info()
        return self

    # MPAS mesh-file variables read by from_netcdf: (reference name, file
    # name, kind); "index" variables are 1-based in the file, 0 = none
    NETCDF_VARIABLES = (
        ("nedges_on_cell", "nEdgesOnCell", "count"), ("edges_on_cell", "edgesOnCell", "index"),
        ("vertices_on_cell", "verticesOnCell", "index"),
        ("nedges_on_edge", "nEdgesOnEdge", "count"), ("edges_on_edge", "edgesOnEdge", "index"),
        ("weights_on_edge", "weightsOnEdge", "real"), ("cells_on_edge", "cellsOnEdge", "index"),
        ("vertices_on_edge", "verticesOnEdge", "index"),
        ("edges_on_vertex", "edgesOnVertex", "index"), ("cells_on_vertex", "cellsOnVertex", "index"),
        ("area_cell", "areaCell", "real"), ("dc_edge", "dcEdge", "real"),
        ("dv_edge", "dvEdge", "real"), ("area_triangle", "areaTriangle", "real"),
        ("kiteareas_on_vertex", "kiteAreasOnVertex", "real"),
        ("x_cell", "xCell", "real"), ("y_cell", "yCell", "real"), ("x_edge", "xEdge", "real"),
        ("y_edge", "yEdge", "real"), ("angle_edge", "angleEdge", "real"),
        ("x_vertex", "xVertex", "real"), ("y_vertex", "yVertex", "real"),
        ("mesh_scaling_del2", "meshScalingDel2", "real"),
        ("mesh_scaling_del4", "meshScalingDel4", "real"),
        ("max_level_edge_top", "maxLevelEdgeTop", "count"),
    )

    @classmethod
    def from_netcdf(cls, path: str, nlayers: Optional[int] = None, ordering: str = "hilbert"):
        """Reads an MPAS mesh file (NetCDF classic or 64-bit-offset format;
        NetCDF-4/HDF5 files need converting first, e.g. `nccopy -k 2`).

        Variables follow the MPAS mesh specification (nEdgesOnCell,
        edgesOnCell, ..., kiteAreasOnVertex, weightsOnEdge, 1-based indices);
        meshScalingDel2/Del4 and maxLevelEdgeTop are used when present.
        nlayers defaults to the file's nVertLevels dimension, else 1.  The
        reference builds meshes in memory only (mpas_mesh.hpp:7-71); this is
        the file route to the same MPASMesh fields."""
        from scipy.io import netcdf_file

        arrays = {}
        with netcdf_file(path, "r", mmap=False) as f:
            var = f.variables
            for name, ncname, kind in cls.NETCDF_VARIABLES:
                if ncname not in var:
                    continue
                x = np.array(var[ncname].data)
                if kind == "real":
                    arrays[name] = x.astype(np.float64)
                else:
                    x = x.astype(np.int32)
                    arrays[name] = x - 1 if kind == "index" else x
            if nlayers is None:
                nlayers = int(f.dimensions.get("nVertLevels") or 1)
        missing = [nc for name, nc, kind in cls.NETCDF_VARIABLES[:15] if name not in arrays]
        if missing:
            raise OmegaError(f"{path}: MPAS mesh variables missing: {', '.join(missing)}")
        return cls.from_arrays(arrays, nlayers, ordering)


class PlanarHexagonalMesh(MPASMesh):
    """PlanarHexagonalMesh(nx, ny[, dc], nlayers=1)
    (src/mesh/planar_hexagonal_mesh.hpp:17-18); dc defaults to 1/nx."""

    def __init__(self, nx: int, ny: int, dc: Optional[float] = None, nlayers: int = 1,
                 ordering: str = "hilbert"):
        if dc is None:
            dc = 1.0 / nx
        h = C.c_void_p()
        check(lib().omega_mesh_create_planar_hex(nx, ny, float(dc), nlayers, ORDER[ordering],
                                                 C.byref(h)))
        self._h = h
        self._init_info()


@dataclasses.dataclass
class ShallowWaterParams:
    """src/model/shallow_water_params.hpp:7-18 (same field order)."""

    f0: float = 0.0
    grav: float = 9.81
    drag_coeff: float = 0.0
    visc_del2: float = 0.0
    visc_del4: float = 0.0
    eddy_diff2: float = 0.0
    eddy_diff4: float = 0.0
    ntracers: int = 0
    disable_h_tendency: bool = False
    disable_vn_tendency: bool = False

    def to_c(self) -> Params:
        return Params(self.f0, self.grav, self.drag_coeff, self.visc_del2, self.visc_del4,
                      self.eddy_diff2, self.eddy_diff4, int(self.ntracers),
                      int(self.disable_h_tendency), int(self.disable_vn_tendency))


class ShallowWaterState:
    """ShallowWaterState(mesh, ntracers) (src/model/shallow_water_state.cpp:7-14)."""

    def __init__(self, mesh: MPASMesh, ntracers):
        if isinstance(ntracers, ShallowWaterParams):
            ntracers = ntracers.ntracers
        self.mesh = mesh
        self.ntracers = int(ntracers)
        h = C.c_void_p()
        check(lib().omega_state_create(mesh._h, self.ntracers, C.byref(h)))
        self._h = h

    def shapes(self):
        m = self.mesh
        return ((m.ncells, m.nlayers), (m.nedges, m.nlayers), (self.ntracers, m.ncells, m.nlayers))

    def upload(self, h=None, vn=None, tr=None, stream=None):
        def prep(a, shp):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            assert a.shape == shp, f"shape {a.shape} != {shp}"
            return a
        sh = self.shapes()
        h, vn, tr = prep(h, sh[0]), prep(vn, sh[1]), prep(tr, sh[2]) if self.ntracers else None
        check(lib().omega_state_upload(self._h, _dp(h), _dp(vn), _dp(tr), _stream(stream)))

    def download(self, stream=None):
        sh = self.shapes()
        h, vn, tr = np.empty(sh[0]), np.empty(sh[1]), np.empty(sh[2])
        check(lib().omega_state_download(self._h, _dp(h), _dp(vn), _dp(tr) if self.ntracers else None,
                                         _stream(stream)))
        return h, vn, tr

    def device_ptrs(self):
        a, b, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().omega_state_device_ptrs(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def copy_from(self, other: "ShallowWaterState", stream=None):
        check(lib().omega_state_copy(self._h, other._h, _stream(stream)))

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.omega_state_destroy(self._h)
            self._h = None


class ShallowWaterModel:
    """ShallowWaterModel(mesh, params) (src/model/shallow_water_model.cpp:5-41).

    Subclasses may override additional_tendency(h_tend, vn_tend, h, vn, t,
    stream) (device pointers, device numbering), the analogue of the
    reference's virtual hook (shallow_water_model.hpp:42-44)."""

    def __init__(self, mesh: MPASMesh, params: ShallowWaterParams):
        self.mesh = mesh
        self.params = params
        h = C.c_void_p()
        self._cparams = params.to_c()
        self._create(h)
        self._h = h
        self._cb = None
        if type(self).additional_tendency is not ShallowWaterModel.additional_tendency:
            def tramp(user, ht, vt, hh, vn, t, stream):
                try:
                    self.additional_tendency(ht, vt, hh, vn, t, stream)
                    return 0
                except Exception as exc:  # reported through the status code
                    self._hook_error = exc
                    return 1
            self._cb = ADDFN(tramp)
            check(lib().omega_model_set_additional_tendency(self._h, self._cb, None))

    def _create(self, h):
        check(lib().omega_model_create(self.mesh._h, C.byref(self._cparams), C.byref(h)))

    def additional_tendency(self, h_tend, vn_tend, h, vn, t, stream):
        pass

    def compute_tendency(self, tend: ShallowWaterState, state: ShallowWaterState, t: float = 0.0,
                         add_mode: int = ADD_REPLACE, stream=None):
        check(lib().omega_model_compute_tendency(self._h, tend._h, state._h, float(t), int(add_mode),
                                                 _stream(stream)))

    def compute_tendency_noalloc(self, tend: ShallowWaterState, state: ShallowWaterState,
                                 t: float = 0.0, add_mode: int = ADD_REPLACE, stream=None):
        """compute_tendency with its diagnostics in the library's static scratch
        array (no allocation; 64 MiB per GPU)."""
        check(lib().omega_model_compute_tendency_noalloc(self._h, tend._h, state._h, float(t),
                                                         int(add_mode), _stream(stream)))

    def _field_tendency(self, field, tend, state, add_mode, stream):
        check(lib().omega_model_compute_field_tendency(self._h, field, tend._h, state._h,
                                                       int(add_mode), _stream(stream)))

    def compute_h_tendency(self, tend, state, add_mode: int = ADD_REPLACE, stream=None):
        """compute_h_tendency (shallow_water_model.cpp:43-86): only tend's h."""
        self._field_tendency(0, tend, state, add_mode, stream)

    def compute_vn_tendency(self, tend, state, add_mode: int = ADD_REPLACE, stream=None):
        """compute_vn_tendency (shallow_water_model.cpp:88-219): only tend's vn."""
        self._field_tendency(1, tend, state, add_mode, stream)

    def compute_tr_tendency(self, tend, state, add_mode: int = ADD_REPLACE, stream=None):
        """compute_tr_tendency (shallow_water_model.cpp:221-311): only tend's tracers."""
        self._field_tendency(2, tend, state, add_mode, stream)

    def compute_aux(self, h, vn=None, tr=None, stream=None):
        """ShallowWaterAuxiliaryState::compute (shallow_water_auxstate.cpp:12-50)
        from device arrays (pointers as ints, device numbering), or from a
        ShallowWaterState passed as h."""
        if isinstance(h, ShallowWaterState):
            h, vn, tr = h.device_ptrs()
        check(lib().omega_model_compute_aux(self._h, C.c_void_p(h), C.c_void_p(vn),
                                            C.c_void_p(tr or 0), _stream(stream)))

    def compute_h_tendency_dev(self, h_tend, h, vn, add_mode: int = ADD_REPLACE, stream=None):
        """compute_h_tendency(h_tend_cell, h_cell, vn_edge, add_mode)
        (shallow_water_model.hpp:35-36) on device pointers; reads the
        auxiliary state of the last compute_aux / tendency / step."""
        check(lib().omega_model_compute_h_tendency(self._h, C.c_void_p(h_tend), C.c_void_p(h),
                                                   C.c_void_p(vn), int(add_mode),
                                                   _stream(stream)))

    def compute_vn_tendency_dev(self, vn_tend, h, vn, add_mode: int = ADD_REPLACE, stream=None):
        """compute_vn_tendency(vn_tend_edge, h_cell, vn_edge, add_mode)
        (shallow_water_model.hpp:37-38) on device pointers."""
        check(lib().omega_model_compute_vn_tendency(self._h, C.c_void_p(vn_tend), C.c_void_p(h),
                                                    C.c_void_p(vn), int(add_mode),
                                                    _stream(stream)))

    def compute_tr_tendency_dev(self, tr_tend, tr, vn, add_mode: int = ADD_REPLACE, stream=None):
        """compute_tr_tendency(tr_tend_cell, tr_cell, vn_edge, add_mode)
        (shallow_water_model.hpp:39-40) on device pointers."""
        check(lib().omega_model_compute_tr_tendency(self._h, C.c_void_p(tr_tend),
                                                    C.c_void_p(tr or 0), C.c_void_p(vn),
                                                    int(add_mode), _stream(stream)))

    def aux_array(self, name: str, stream=None) -> np.ndarray:
        """Host copy (original numbering) of an auxiliary-state array:
        rvort_vertex, inv_thick_vertex, ke_cell, vel_div_cell, tracer_del2_cell."""
        m = self.mesh
        L = m.nlayers
        shape = {"rvort_vertex": (m.nvertices, L), "inv_thick_vertex": (m.nvertices, L),
                 "ke_cell": (m.ncells, L), "vel_div_cell": (m.ncells, L),
                 "tracer_del2_cell": (self.params.ntracers, m.ncells, L)}.get(name)
        if shape is None:
            raise OmegaError(f"unknown auxiliary array {name}")
        out = np.empty(shape)
        check(lib().omega_model_aux_download(self._h, name.encode(), _dp(out), _stream(stream)))
        return out

    @property
    def terms(self) -> int:
        m = C.c_int()
        check(lib().omega_model_get_terms(self._h, C.byref(m)))
        return m.value

    @terms.setter
    def terms(self, mask: int):
        check(lib().omega_model_set_terms(self._h, int(mask)))

    def set_launch(self, rows_per_block=0, levels_per_thread=0, xcd_swizzle=-1,
                   levels_per_block=-1):
        check(lib().omega_model_set_launch(self._h, rows_per_block, levels_per_thread, xcd_swizzle,
                                           levels_per_block))

    def get_launch(self):
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().omega_model_get_launch(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return a.value, b.value, c.value, d.value

    def autotune(self, state: "ShallowWaterState", reps: int = 3, stream=None) -> float:
        """Times the launch shapes on tendency evaluations of state and keeps the
        fastest; returns its milliseconds per evaluation."""
        ms = C.c_double()
        check(lib().omega_model_autotune(self._h, state._h, int(reps), C.byref(ms),
                                         _stream(stream)))
        return ms.value

    def set_variant(self, variant: int):
        """0: pattern-specialised kernels when possible, 1: general kernels."""
        check(lib().omega_model_set_variant(self._h, int(variant)))

    def kernel_flags(self) -> int:
        """bit 0: coefficient tables uniform per stencil slot (shared by a
        wavefront's two cells); bit 1: 32-bit row offsets; bit 2: the
        tendency VJP's slot-uniform form (after the first adjoint call);
        bit 3: the pair tendency kernel compiled for the model's term set;
        bit 4: a step adjoint's stage recomputation writes the rows its
        tendency VJPs linearise about."""
        v = C.c_int()
        check(lib().omega_model_kernel_flags(self._h, C.byref(v)))
        return v.value

    def set_kernel_forms(self, forms: int = 15):
        """Allow (bits set) or forbid the shared-coefficient (bit 0),
        32-bit-offset (bit 1) and compiled-in term-set (bit 2) forms of the
        pair kernels, and a step adjoint's stage recomputation writing the
        rows its tendency VJPs linearise about (bit 3)."""
        check(lib().omega_model_set_kernel_forms(self._h, int(forms)))

    def kernel_variant(self) -> int:
        """6 linear model, 2 pattern-specialised, 1 general kernels."""
        v = C.c_int()
        check(lib().omega_model_kernel_variant(self._h, C.byref(v)))
        return v.value

    def set_f_vertex(self, f_vertex: np.ndarray):
        f = np.ascontiguousarray(f_vertex, dtype=np.float64)
        check(lib().omega_model_set_f_vertex(self._h, _dp(f)))

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.omega_model_destroy(self._h)
            self._h = None


class LinearShallowWaterModel(ShallowWaterModel):
    """LinearShallowWaterModel(mesh, params, h0)
    (src/shallow_water_hierarchical.cpp:676-773): the equations linearised
    about a resting layer of thickness h0, h_t = -h0 div(vn),
    vn_t = f0 * (TRiSK tangential velocity) - g grad(h); no tracers
    (params.ntracers must be 0).  Steppers, hooks, increment mode and the
    integrals work as for ShallowWaterModel; the adjoint and ensembles
    reject it."""

    def __init__(self, mesh: MPASMesh, params: ShallowWaterParams, h0: float):
        self.h0 = float(h0)
        super().__init__(mesh, params)

    def _create(self, h):
        check(lib().omega_model_create_linear(self.mesh._h, C.byref(self._cparams), self.h0,
                                              C.byref(h)))


class _Stepper:
    _kind = 0

    def __init__(self, model: ShallowWaterModel, nstages: int = 5):
        self.model = model
        h = C.c_void_p()
        check(lib().omega_stepper_create(model._h, self._kind, int(nstages), C.byref(h)))
        self._h = h

    def do_step(self, t: float, dt: float, state: ShallowWaterState, stream=None):
        check(lib().omega_stepper_do_step(self._h, float(t), float(dt), state._h, _stream(stream)))

    def run(self, t0: float, dt: float, nsteps: int, state: ShallowWaterState,
            t_offset_steps: int = 0, stream=None):
        check(lib().omega_stepper_run(self._h, float(t0), float(dt), int(nsteps), int(t_offset_steps),
                                      state._h, _stream(stream)))

    def set_graph(self, enable: bool = True):
        check(lib().omega_stepper_set_graph(self._h, int(bool(enable))))

    def coeffs(self):
        """(m_rka, m_rkb, m_rkc), nstages values each."""
        n = C.c_int()
        check(lib().omega_stepper_coeffs(self._h, C.byref(n), None, None, None))
        a, b, c = (np.zeros(n.value) for _ in range(3))
        check(lib().omega_stepper_coeffs(self._h, C.byref(n), _dp(a), _dp(b), _dp(c)))
        return a, b, c

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.omega_stepper_destroy(self._h)
            self._h = None


class RK4Stepper(_Stepper):
    """RK4Stepper(model) (src/time_integration/rk4.cpp)."""

    _kind = 0

    def __init__(self, model: ShallowWaterModel):
        super().__init__(model, 4)


class LSRKStepper(_Stepper):
    """LSRKStepper(model, nstages=5) (src/time_integration/lsrk.cpp)."""

    _kind = 1


# ---------------------------------------------------------------------------
# Discrete adjoint (no reference counterpart; differentiates the operators of
# src/model/shallow_water_model.cpp and src/time_integration/{rk4,lsrk}.cpp)
# ---------------------------------------------------------------------------
PARAM_NAMES = ("visc_del2", "visc_del4", "eddy_diff2", "eddy_diff4")


def _pgrad_out(param_grad):
    if param_grad is None:
        return None, None
    buf = (C.c_double * 4)(*[float(x) for x in param_grad])
    return buf, C.cast(buf, C.POINTER(C.c_double))


def tendency_vjp(model: ShallowWaterModel, state: ShallowWaterState, tend_cot: ShallowWaterState,
                 state_cot: ShallowWaterState, stream=None) -> np.ndarray:
    """state_cot += J(state)^T tend_cot; returns d<tend_cot, F>/d(coefficients)
    in the order of PARAM_NAMES."""
    buf, ptr = _pgrad_out([0.0] * 4)
    check(lib().omega_model_tendency_vjp(model._h, state._h, tend_cot._h, state_cot._h, ptr,
                                         _stream(stream)))
    return np.array(buf[:])


def step_vjp(stepper: "_Stepper", t: float, dt: float, state: ShallowWaterState,
             cot: ShallowWaterState, stream=None, pgrad: bool = True):
    """cot <- (dy_{n+1}/dy_n)^T cot for one step from state; returns the
    coefficient gradient contribution (read back to the host, which waits
    for the stream), or None with pgrad=False (the call then returns
    without waiting)."""
    if not pgrad:
        check(lib().omega_stepper_step_vjp(stepper._h, float(t), float(dt), state._h, cot._h,
                                           None, _stream(stream)))
        return None
    buf, ptr = _pgrad_out([0.0] * 4)
    check(lib().omega_stepper_step_vjp(stepper._h, float(t), float(dt), state._h, cot._h, ptr,
                                       _stream(stream)))
    return np.array(buf[:])


def revolve_cost(nsteps: int, n_snapshots: int) -> int:
    """Minimal forward steps of a binomial reversal schedule."""
    return int(lib().omega_revolve_cost(int(nsteps), int(n_snapshots)))


class Tape:
    """Binomial checkpointing over nsteps steps of a stepper."""

    def __init__(self, stepper: "_Stepper", nsteps: int, n_snapshots: int, byte_cap: int = 0):
        self.stepper = stepper
        h = C.c_void_p()
        check(lib().omega_tape_create(stepper._h, int(nsteps), int(n_snapshots), int(byte_cap),
                                      C.byref(h)))
        self._h = h

    def forward(self, t0: float, dt: float, state: ShallowWaterState, stream=None):
        check(lib().omega_tape_forward(self._h, float(t0), float(dt), state._h, _stream(stream)))

    def reverse(self, cot: ShallowWaterState, stream=None) -> np.ndarray:
        buf, ptr = _pgrad_out([0.0] * 4)
        check(lib().omega_tape_reverse(self._h, cot._h, ptr, _stream(stream)))
        return np.array(buf[:])

    def adjoint_run(self, t0: float, nsteps: int, obs, grads, weights=(1.0, 0.0, 1.0),
                    n_snapshots: int = 10, byte_cap: int = 0, stream=None):
        """Ensemble-mean misfit J = (1/M) sum_m J_m after nsteps steps and its
        gradient: grads[m] = dJ/d(member m's initial state); returns (J,
        gradient w.r.t. the shared coefficients, in the order of
        PARAM_NAMES).  Members must share one mesh, layer and tracer counts,
        parameters and dt; they are left unchanged."""
        M = len(self._members)
        if len(obs) != M or len(grads) != M:
            raise ValueError("one observation and one gradient state per member")
        o = (C.c_void_p * M)(*[x._h.value if isinstance(x._h, C.c_void_p) else x._h for x in obs])
        g = (C.c_void_p * M)(*[x._h.value if isinstance(x._h, C.c_void_p) else x._h for x in grads])
        w = (C.c_double * 3)(*[float(x) for x in weights])
        J = C.c_double()
        buf, ptr = _pgrad_out([0.0] * 4)
        check(lib().omega_batch_adjoint_run(self._h, float(t0), int(nsteps), int(n_snapshots),
                                            int(byte_cap), o, w, C.byref(J), g, ptr,
                                            _stream(stream)))
        return J.value, np.array(buf[:])

    def info(self):
        n, fs, nb = C.c_int(), C.c_int64(), C.c_int64()
        check(lib().omega_tape_info(self._h, C.byref(n), C.byref(fs), C.byref(nb)))
        return {"n_snapshots": n.value, "forward_steps": fs.value, "snapshot_bytes": nb.value}

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.omega_tape_destroy(self._h)
            self._h = None


def misfit(model: ShallowWaterModel, state: ShallowWaterState, obs: ShallowWaterState,
           cot: ShallowWaterState, weights=(1.0, 0.0, 1.0), stream=None) -> float:
    """J = sum_f w_f/2 sum_i area_i sum_k (y - obs)^2; cot = dJ/dy."""
    w = (C.c_double * 3)(*[float(x) for x in weights])
    J = C.c_double()
    check(lib().omega_misfit(model._h, state._h, obs._h, w, cot._h, C.byref(J), _stream(stream)))
    return J.value


def adjoint_run(stepper: "_Stepper", t0: float, dt: float, nsteps: int, state: ShallowWaterState,
                obs: ShallowWaterState, grad: ShallowWaterState, weights=(1.0, 0.0, 1.0),
                n_snapshots: int = 10, byte_cap: int = 0, stream=None):
    """Misfit after nsteps steps from state and its gradient: grad = dJ/dy_0;
    returns (J, coefficient gradient)."""
    w = (C.c_double * 3)(*[float(x) for x in weights])
    J = C.c_double()
    buf, ptr = _pgrad_out([0.0] * 4)
    check(lib().omega_adjoint_run(stepper._h, float(t0), float(dt), int(nsteps), int(n_snapshots),
                                  int(byte_cap), state._h, obs._h, w, C.byref(J), grad._h, ptr,
                                  _stream(stream)))
    return J.value, np.array(buf[:])


# ---------------------------------------------------------------------------
# Ensembles (no reference counterpart)
# ---------------------------------------------------------------------------
class Batch:
    """Ensemble of independent (model, state, dt) members advanced together,
    one launch per stage phase for all members."""

    def __init__(self, kind: str = "lsrk", nstages: int = 5):
        h = C.c_void_p()
        check(lib().omega_batch_create(0 if kind == "rk4" else 1, int(nstages), C.byref(h)))
        self._h = h
        self._members = []  # keeps models and states alive

    def add(self, model: ShallowWaterModel, state: ShallowWaterState, dt: float) -> int:
        idx = C.c_int()
        check(lib().omega_batch_add(self._h, model._h, state._h, float(dt), C.byref(idx)))
        self._members.append((model, state))
        return idx.value

    def step(self, stream=None):
        check(lib().omega_batch_step(self._h, _stream(stream)))

    def run(self, nsteps: int, stream=None):
        check(lib().omega_batch_run(self._h, int(nsteps), _stream(stream)))

    def set_graph(self, enable: bool = True):
        check(lib().omega_batch_set_graph(self._h, int(bool(enable))))

    def set_launch(self, threads: int = 256, levels_per_thread: int = 2):
        check(lib().omega_batch_set_launch(self._h, int(threads), int(levels_per_thread)))

    def get_launch(self):
        t, n = C.c_int(), C.c_int()
        check(lib().omega_batch_get_launch(self._h, C.byref(t), C.byref(n)))
        return t.value, n.value

    def set_persistent(self, blocks_per_cu: int = 0):
        """0: one block per table entry; n > 0: a persistent grid of n blocks per CU."""
        check(lib().omega_batch_set_persistent(self._h, int(blocks_per_cu)))

    def get_persistent(self) -> int:
        v = C.c_int()
        check(lib().omega_batch_get_persistent(self._h, C.byref(v)))
        return v.value

    def autotune(self, reps: int = 3, stream=None) -> float:
        ms = C.c_double()
        check(lib().omega_batch_autotune(self._h, int(reps), C.byref(ms), _stream(stream)))
        return ms.value

    def adjoint_run(self, t0: float, nsteps: int, obs, grads, weights=(1.0, 0.0, 1.0),
                    n_snapshots: int = 10, byte_cap: int = 0, stream=None):
        """Ensemble-mean misfit J = (1/M) sum_m J_m after nsteps steps and its
        gradient: grads[m] = dJ/d(member m's initial state); returns (J,
        gradient w.r.t. the shared coefficients, in the order of
        PARAM_NAMES).  Members must share one mesh, layer and tracer counts,
        parameters and dt; they are left unchanged."""
        M = len(self._members)
        if len(obs) != M or len(grads) != M:
            raise ValueError("one observation and one gradient state per member")
        o = (C.c_void_p * M)(*[x._h.value if isinstance(x._h, C.c_void_p) else x._h for x in obs])
        g = (C.c_void_p * M)(*[x._h.value if isinstance(x._h, C.c_void_p) else x._h for x in grads])
        w = (C.c_double * 3)(*[float(x) for x in weights])
        J = C.c_double()
        buf, ptr = _pgrad_out([0.0] * 4)
        check(lib().omega_batch_adjoint_run(self._h, float(t0), int(nsteps), int(n_snapshots),
                                            int(byte_cap), o, w, C.byref(J), g, ptr,
                                            _stream(stream)))
        return J.value, np.array(buf[:])

    def info(self):
        n, b, w = C.c_int(), C.c_int64(), C.c_int64()
        check(lib().omega_batch_info(self._h, C.byref(n), C.byref(b), C.byref(w)))
        return {"members": n.value, "blocks_per_launch": b.value, "cell_layers": w.value}

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib is not None:
            _lib.omega_batch_destroy(self._h)
            self._h = None


def _integral(fn, state, model, stream=None):
    out = C.c_double()
    check(fn(model._h, state._h, C.byref(out), _stream(stream)))
    return out.value


def energy_integral(state: ShallowWaterState, model: ShallowWaterModel, stream=None) -> float:
    """src/model/shallow_water_diagnostics.cpp:5-38."""
    return _integral(lib().omega_energy_integral, state, model, stream)


def mass_integral(state: ShallowWaterState, model: ShallowWaterModel, stream=None) -> float:
    """src/model/shallow_water_diagnostics.cpp:40-58."""
    return _integral(lib().omega_mass_integral, state, model, stream)


def circulation_integral(state: ShallowWaterState, model: ShallowWaterModel, stream=None) -> float:
    """src/model/shallow_water_diagnostics.cpp:60-89."""
    return _integral(lib().omega_circulation_integral, state, model, stream)


def synchronize(stream=None):
    check(lib().omega_synchronize(_stream(stream)))


# timing helpers ------------------------------------------------------------
def timing_enable(on: bool = True):
    check(lib().omega_timing_enable(int(on)))


def timing_reset():
    check(lib().omega_timing_reset())


def timing_read(name: str):
    ms, n = C.c_double(), C.c_int64()
    check(lib().omega_timing_read(name.encode(), C.byref(ms), C.byref(n)))
    return ms.value, n.value
