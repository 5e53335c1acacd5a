"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, always as the checker.  The product path (omega_playground_amd)
never imports this module.

The oracle restates the reference's arithmetic (see swe_oracle.cpp for the
file:line citations).  State arrays are caller-owned numpy float64 arrays in
the reference's original entity numbering and LayoutRight shape:
h[ncells, L], vn[nedges, L], tr[ntracers, ncells, L].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build (the mutation check's planted-mistake builds)
_LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle_swe.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++)."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class Params(C.Structure):
    # src/model/shallow_water_params.hpp:7-18 (same field order)
    _fields_ = [
        ("f0", C.c_double),
        ("grav", C.c_double),
        ("drag_coeff", C.c_double),
        ("visc_del2", C.c_double),
        ("visc_del4", C.c_double),
        ("eddy_diff2", C.c_double),
        ("eddy_diff4", C.c_double),
        ("ntracers", C.c_int),
        ("disable_h_tendency", C.c_int),
        ("disable_vn_tendency", C.c_int),
    ]


ADDFN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        lp = C.POINTER(C.c_long)
        L.orc_mesh_planar_hex.restype = vp
        L.orc_mesh_planar_hex.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int]
        L.orc_mesh_free.argtypes = [vp]
        L.orc_mesh_info.argtypes = [vp, ip, dp]
        L.orc_mesh_array.restype = vp
        L.orc_mesh_array.argtypes = [vp, C.c_char_p, ip, lp, lp]
        L.orc_mesh_finalize.argtypes = [vp]
        L.orc_mesh_alloc.restype = vp
        L.orc_mesh_alloc.argtypes = [C.c_int] * 4
        L.orc_model_create.restype = vp
        L.orc_model_create.argtypes = [vp, C.POINTER(Params)]
        L.orc_model_free.argtypes = [vp]
        L.orc_model_set_linear.argtypes = [vp, C.c_double]
        L.orc_model_array.restype = vp
        L.orc_model_array.argtypes = [vp, C.c_char_p, ip, lp, lp]
        L.orc_model_set_additional.argtypes = [vp, ADDFN, vp]
        L.orc_model_set_terms.argtypes = [vp, C.c_int]
        L.orc_model_get_terms.argtypes = [vp]
        L.orc_model_get_terms.restype = C.c_int
        L.orc_aux_compute.argtypes = [vp, dp, dp, dp]
        L.orc_compute_tendency.argtypes = [vp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_int]
        L.orc_compute_field_tendency.argtypes = [vp, C.c_int, dp, dp, dp, C.c_int]
        L.orc_stepper_create.restype = vp
        L.orc_stepper_create.argtypes = [vp, C.c_int, C.c_int]
        L.orc_stepper_free.argtypes = [vp]
        L.orc_stepper_coeffs.argtypes = [vp, dp, dp, dp]
        L.orc_stepper_nstages.argtypes = [vp]
        L.orc_stepper_nstages.restype = C.c_int
        L.orc_do_steps.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp, dp]
        for fn in ("orc_energy",):
            getattr(L, fn).restype = C.c_double
            getattr(L, fn).argtypes = [vp, dp, dp]
        L.orc_mass.restype = C.c_double
        L.orc_mass.argtypes = [vp, dp]
        L.orc_circulation.restype = C.c_double
        L.orc_circulation.argtypes = [vp, dp]
        L.orc_set_num_threads.argtypes = [C.c_int]
        L.orc_get_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def max_threads() -> int:
    return lib().orc_get_max_threads()


def _view(ptr, dtype, n0, n1):
    if not ptr:
        raise KeyError("no such array")
    ct = C.c_int if dtype == 0 else C.c_double
    npd = np.int32 if dtype == 0 else np.float64
    count = n0 * n1
    buf = (ct * count).from_address(ptr)
    a = np.frombuffer(buf, dtype=npd, count=count)
    return a.reshape(n0, n1) if n1 != 1 else a


class Mesh:
    """Planar hexagonal mesh, reference numbering
    (src/mesh/planar_hexagonal_mesh.cpp:8-398)."""

    def __init__(self, nx: int, ny: int, dc: float | None = None, nlayers: int = 1):
        if dc is None:
            dc = 1.0 / nx  # planar_hexagonal_mesh.cpp:5-6
        self._h = lib().orc_mesh_planar_hex(nx, ny, dc, nlayers)
        ints = (C.c_int * 6)()
        reals = (C.c_double * 3)()
        lib().orc_mesh_info(self._h, ints, reals)
        (self.nx, self.ny, self.nlayers, self.ncells, self.nedges, self.nvertices) = list(ints)
        self.dc, self.period_x, self.period_y = list(reals)

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        dt, n0, n1 = C.c_int(), C.c_long(), C.c_long()
        p = lib().orc_mesh_array(self._h, name.encode(), C.byref(dt), C.byref(n0), C.byref(n1))
        if not p:
            raise AttributeError(name)
        return _view(p, dt.value, n0.value, n1.value)

    @classmethod
    def from_arrays(cls, arrays: dict, nlayers: int) -> "Mesh":
        """A general MPASMesh filled from arrays (reference names, 0-based,
        rows padded to maxedges / 2*maxedges columns), then finalize_mesh()
        (src/mesh/mpas_mesh.hpp:7-71, mpas_mesh.cpp:5-57)."""
        self = cls.__new__(cls)
        nc = len(arrays["nedges_on_cell"])
        ne = len(arrays["nedges_on_edge"])
        nv = len(arrays["edges_on_vertex"])
        self._h = lib().orc_mesh_alloc(nc, ne, nv, nlayers)
        self.nx = self.ny = 0
        self.nlayers, self.ncells, self.nedges, self.nvertices = nlayers, nc, ne, nv
        self.dc = self.period_x = self.period_y = 0.0
        for name, a in arrays.items():
            if name in ("mesh_scaling_del2", "mesh_scaling_del4", "max_level_edge_top"):
                continue
            dst = getattr(self, name)
            a = np.asarray(a)
            if dst.ndim == 2 and a.ndim == 2 and a.shape[1] < dst.shape[1]:
                dst[:, :a.shape[1]] = a
            else:
                dst[...] = a
        self.finalize()
        # optional per-edge scalings and level bounds replace finalize_mesh's
        # unit defaults; edge_mask follows max_level_edge_top
        for name in ("mesh_scaling_del2", "mesh_scaling_del4"):
            if arrays.get(name) is not None:
                getattr(self, name)[:] = arrays[name]
        if arrays.get("max_level_edge_top") is not None:
            top = np.asarray(arrays["max_level_edge_top"])
            self.max_level_edge_top[:] = top
            self.edge_mask[:] = (np.arange(nlayers)[None, :] < top[:, None]).astype(np.float64)
        return self

    def finalize(self):
        lib().orc_mesh_finalize(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_mesh_free(self._h)
            self._h = None


PV, KE, SSH, VDEL2, VDEL4, TDEL2, TDEL4, DRAG = 1, 2, 4, 8, 16, 32, 64, 128


class Model:
    """ShallowWaterModel restatement (src/model/shallow_water_model.cpp)."""

    def __init__(self, mesh: Mesh, f0=0.0, grav=9.81, drag_coeff=0.0, visc_del2=0.0,
                 visc_del4=0.0, eddy_diff2=0.0, eddy_diff4=0.0, ntracers=0,
                 disable_h_tendency=False, disable_vn_tendency=False, linear_h0=None):
        self.mesh = mesh
        self.params = Params(f0, grav, drag_coeff, visc_del2, visc_del4, eddy_diff2, eddy_diff4,
                             int(ntracers), int(disable_h_tendency), int(disable_vn_tendency))
        self._h = lib().orc_model_create(mesh._h, C.byref(self.params))
        self._cb = None
        self.ntracers = int(ntracers)
        if linear_h0 is not None:
            # LinearShallowWaterModel (src/shallow_water_hierarchical.cpp:676-773)
            if self.ntracers:
                raise ValueError("the linear model carries no tracers")
            lib().orc_model_set_linear(self._h, float(linear_h0))

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        dt, n0, n1 = C.c_int(), C.c_long(), C.c_long()
        p = lib().orc_model_array(self._h, name.encode(), C.byref(dt), C.byref(n0), C.byref(n1))
        if not p:
            raise AttributeError(name)
        a = _view(p, dt.value, n0.value, n1.value)
        # level-resolved arrays stay [entity, level] even with one level
        return a if name in ("f_vertex", "f_edge") else a.reshape(n0.value, n1.value)

    @property
    def terms(self) -> int:
        return lib().orc_model_get_terms(self._h)

    @terms.setter
    def terms(self, mask: int) -> None:
        lib().orc_model_set_terms(self._h, int(mask))

    def set_additional_tendency(self, fn):
        """fn(h_tend, vn_tend, h, vn, t) with numpy views, modifies tendencies in place."""
        m = self.mesh
        L = m.nlayers

        def tramp(user, ht, vt, h, vn, t):
            a = np.ctypeslib.as_array(ht, shape=(m.ncells, L))
            b = np.ctypeslib.as_array(vt, shape=(m.nedges, L))
            c = np.ctypeslib.as_array(h, shape=(m.ncells, L))
            d = np.ctypeslib.as_array(vn, shape=(m.nedges, L))
            fn(a, b, c, d, t)

        self._cb = ADDFN(tramp) if fn is not None else None
        lib().orc_model_set_additional(self._h, self._cb if fn is not None else ADDFN(), None)

    def new_state(self):
        m = self.mesh
        L = m.nlayers
        return (np.zeros((m.ncells, L)), np.zeros((m.nedges, L)),
                np.zeros((self.ntracers, m.ncells, L)))

    def aux_compute(self, h, vn, tr):
        lib().orc_aux_compute(self._h, _dptr(h), _dptr(vn), _dptr(tr))

    def compute_field_tendency(self, field, tend, a, vn, increment=False):
        """compute_h_tendency (field 0, a = h), compute_vn_tendency (1, a = h)
        or compute_tr_tendency (2, a = tr), reading the auxiliary state of
        the last aux_compute."""
        lib().orc_compute_field_tendency(self._h, int(field), _dptr(tend), _dptr(a), _dptr(vn),
                                         int(bool(increment)))

    def compute_tendency(self, tend, state, t=0.0, increment=False):
        ht, vt, tt = tend
        h, vn, tr = state
        lib().orc_compute_tendency(self._h, _dptr(ht), _dptr(vt), _dptr(tt), _dptr(h), _dptr(vn),
                                   _dptr(tr), float(t), int(bool(increment)))

    def energy(self, h, vn):
        return lib().orc_energy(self._h, _dptr(h), _dptr(vn))

    def mass(self, h):
        return lib().orc_mass(self._h, _dptr(h))

    def circulation(self, vn):
        return lib().orc_circulation(self._h, _dptr(vn))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_model_free(self._h)
            self._h = None


class Stepper:
    """RK4Stepper (kind='rk4', src/time_integration/rk4.cpp) or
    LSRKStepper (kind='lsrk', src/time_integration/lsrk.cpp)."""

    def __init__(self, model: Model, kind: str = "rk4", nstages: int = 5):
        self.model = model
        self._h = lib().orc_stepper_create(model._h, 0 if kind == "rk4" else 1, int(nstages))
        self.nstages = lib().orc_stepper_nstages(self._h)

    def coeffs(self):
        a = np.zeros(self.nstages)
        b = np.zeros(self.nstages)
        c = np.zeros(self.nstages)
        lib().orc_stepper_coeffs(self._h, _dptr(a), _dptr(b), _dptr(c))
        return a, b, c

    def do_step(self, t, dt, state):
        self.run(t, dt, 1, state)

    def run(self, t0, dt, nsteps, state, t_from_step_plus_one=False):
        """nsteps steps with t = t0 + step*dt (or (step+1)*dt)."""
        h, vn, tr = state
        lib().orc_do_steps(self._h, float(t0), float(dt), int(nsteps),
                           int(bool(t_from_step_plus_one)), _dptr(h), _dptr(vn), _dptr(tr))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_stepper_free(self._h)
            self._h = None
