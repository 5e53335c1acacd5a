"""General MPAS meshes on the CPU: the icosahedral test mesh (pentagons and
hexagons, 9 or 10 edges-on-edge), its TRiSK weights and its MPAS NetCDF
file, read back through MPASMesh.from_netcdf.

The generator is pinned three ways: its weight routine reproduces the
reference's planar-mesh weights and edges-on-edge order
(planar_hexagonal_mesh.cpp:199-292, through the oracle's restatement),
and on the sphere the oracle shows the identities TRiSK guarantees for a
consistent mesh: the curl of a gradient vanishes at every vertex, the PV
flux term does no work (energy neutral), and thickness advection
conserves mass.  A weight with a dropped sign breaks the energy identity
(checked below)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from omega_playground_amd import swe

import meshgen

PRIMARY = ("nedges_on_cell", "edges_on_cell", "vertices_on_cell", "cells_on_edge",
           "cells_on_vertex", "kiteareas_on_vertex", "area_cell", "dc_edge", "dv_edge")


@pytest.fixture(scope="module")
def ico():
    return meshgen.icosahedral_mesh(2, seed=1)


def test_trisk_weights_reproduce_reference_planar_weights():
    o = po.Mesh(8, 6, 2.0, 3)
    a = {f: np.array(getattr(o, f)) for f in PRIMARY}
    n, eoe, w = meshgen.trisk_weights(a)
    assert np.array_equal(n, o.nedges_on_edge)
    assert np.array_equal(eoe[:, :10], o.edges_on_edge[:, :10])
    assert np.abs(w - o.weights_on_edge).max() <= 1e-15


def test_icosahedral_mesh_shape(ico):
    nc, ne, nv = len(ico["area_cell"]), len(ico["dc_edge"]), len(ico["area_triangle"])
    assert (nc, ne, nv) == (162, 480, 320)
    assert nc - ne + nv == 2
    assert np.array_equal(np.bincount(ico["nedges_on_cell"]), [0, 0, 0, 0, 0, 12, 150])
    assert np.array_equal(np.bincount(ico["nedges_on_edge"], minlength=11)[9:], [60, 420])
    R = 6.371e6
    assert ico["kiteareas_on_vertex"].min() > 0
    assert abs(ico["area_cell"].sum() / (4 * np.pi * R * R) - 1) <= 1e-13
    assert abs(ico["area_triangle"].sum() / (4 * np.pi * R * R) - 1) <= 1e-13
    # first/second cell of an edge both occur with each cell slot
    assert 0.3 < (ico["cells_on_edge"][:, 0] < ico["cells_on_edge"][:, 1]).mean() < 0.7


def _oracle(ico, L, ntr, terms):
    om = po.Mesh.from_arrays(ico, L)
    m = po.Model(om, f0=1e-4, ntracers=ntr)
    m.terms = terms
    return om, m


def test_icosahedral_curl_of_gradient_vanishes(ico):
    L = 3
    om, m = _oracle(ico, L, 0, po.PV)
    phi = np.random.default_rng(1).standard_normal((len(ico["area_cell"]), L))
    coe = ico["cells_on_edge"]
    vn = np.ascontiguousarray((phi[coe[:, 1]] - phi[coe[:, 0]]) / ico["dc_edge"][:, None])
    m.aux_compute(np.full(phi.shape, 1000.0), vn, np.zeros((0,) + phi.shape))
    scale = np.abs(vn).max() / ico["dc_edge"].min()
    assert np.abs(m.rvort_vertex).max() <= 1e-13 * scale


def _pv_work(ico, L):
    om, m = _oracle(ico, L, 1, po.PV)
    h, vn, tr = meshgen.sphere_state(ico, L, 1, 3)
    t = m.new_state()
    m.compute_tendency(t, (h, vn, tr))
    x = (ico["dc_edge"] * ico["dv_edge"])[:, None] * m.flux_h_edge * vn * t[1]
    return x.sum(), np.abs(x).sum()


def test_icosahedral_pv_flux_is_energy_neutral(ico):
    work, scale = _pv_work(ico, 4)
    assert abs(work) <= 1e-14 * scale
    # a weight that loses the sign of its cell's orientation breaks it
    bad = dict(ico)
    w = ico["weights_on_edge"].copy()
    w[:, :4] *= -1
    bad["weights_on_edge"] = w
    work, scale = _pv_work(bad, 4)
    assert abs(work) > 1e-3 * scale


def test_icosahedral_thickness_advection_conserves_mass(ico):
    L = 4
    om, m = _oracle(ico, L, 1, po.PV | po.KE | po.SSH | po.TDEL2 | po.TDEL4)
    h, vn, tr = meshgen.sphere_state(ico, L, 1, 5)
    t = m.new_state()
    m.compute_tendency(t, (h, vn, tr))
    x = ico["area_cell"][:, None] * t[0]
    assert abs(x.sum()) <= 1e-14 * np.abs(x).sum()
    y = ico["area_cell"][:, None] * t[2][0]
    assert abs(y.sum()) <= 1e-13 * np.abs(y).sum()


HOST_FIELDS = ("nedges_on_cell", "edges_on_cell", "vertices_on_cell", "nedges_on_edge",
               "edges_on_edge", "weights_on_edge", "cells_on_edge", "vertices_on_edge",
               "edges_on_vertex", "cells_on_vertex", "area_cell", "dc_edge", "dv_edge",
               "area_triangle", "kiteareas_on_vertex", "edge_sign_on_cell",
               "edge_sign_on_vertex", "kite_index_on_cell", "mesh_scaling_del2",
               "mesh_scaling_del4", "max_level_edge_top")


# entries past a row's count are padding (-1 read from a file)
ROW_COUNT = {"edges_on_cell": "nedges_on_cell", "vertices_on_cell": "nedges_on_cell",
             "edge_sign_on_cell": "nedges_on_cell", "kite_index_on_cell": "nedges_on_cell",
             "edges_on_edge": "nedges_on_edge", "weights_on_edge": "nedges_on_edge"}


def _rows(m, f, ncols=None):
    x = np.array(getattr(m, f))
    if f in ROW_COUNT:
        cnt = np.asarray(getattr(m, ROW_COUNT[f]))
        x = x[:, :ncols] if ncols else x
        x[np.arange(x.shape[1])[None, :] >= cnt[:, None]] = 0
    return x


def _same_mesh(g, h):
    for f in HOST_FIELDS:
        a, b = _rows(g, f), _rows(h, f)
        assert a.shape == b.shape and a.dtype == b.dtype, f
        assert a.tobytes() == b.tobytes(), f
    assert np.array_equal(g.permutation(0), h.permutation(0))
    assert np.array_equal(g.permutation(1), h.permutation(1))


@pytest.mark.parametrize("ordering", ["original", "hilbert"])
def test_netcdf_icosahedral_roundtrip(ico, tmp_path, ordering):
    """Written as an MPAS file (1-based, rows padded with 0) and read back,
    the mesh equals the one built from the arrays, bit for bit."""
    path = str(tmp_path / "ico.nc")
    meshgen.write_mpas_netcdf(path, ico, nlayers=5, sphere_radius=6.371e6)
    g = swe.MPASMesh.from_netcdf(path, ordering=ordering)
    assert g.nlayers == 5 and not g.regular_hex
    h = swe.MPASMesh.from_arrays(ico, 5, ordering)
    _same_mesh(g, h)


def test_netcdf_wider_rows_and_optional_fields(tmp_path):
    """maxEdges = 8 padding, per-edge mesh scalings and level bounds."""
    o = po.Mesh(10, 8, 2.0, 6)
    a = {f: np.array(getattr(o, f)) for f in (
        "nedges_on_cell", "edges_on_cell", "vertices_on_cell", "nedges_on_edge", "edges_on_edge",
        "cells_on_edge", "vertices_on_edge", "edges_on_vertex", "cells_on_vertex", "area_cell",
        "dc_edge", "dv_edge", "weights_on_edge", "area_triangle", "kiteareas_on_vertex",
        "x_cell", "y_cell", "x_edge", "y_edge", "angle_edge", "x_vertex", "y_vertex")}
    rng = np.random.default_rng(2)
    a["mesh_scaling_del2"] = rng.uniform(0.5, 1.5, o.nedges)
    a["mesh_scaling_del4"] = rng.uniform(0.5, 1.5, o.nedges)
    a["max_level_edge_top"] = rng.integers(2, 7, o.nedges).astype(np.int32)
    path = str(tmp_path / "hex.nc")
    meshgen.write_mpas_netcdf(path, a, max_edges=8)
    g = swe.MPASMesh.from_netcdf(path, nlayers=6, ordering="hilbert")
    h = swe.MPASMesh.from_arrays(a, 6, "hilbert")
    for f in HOST_FIELDS:
        y = _rows(h, f)
        x = _rows(g, f, y.shape[1] if y.ndim == 2 else None)
        assert x.tobytes() == y.tobytes(), f
    assert g.regular_hex and h.regular_hex


def test_netcdf_missing_variable_is_an_error(tmp_path):
    from scipy.io import netcdf_file
    path = str(tmp_path / "bad.nc")
    with netcdf_file(path, "w", version=2) as f:
        f.createDimension("nCells", 3)
        f.createVariable("areaCell", "d", ("nCells",))[:] = 1.0
    with pytest.raises(swe.OmegaError, match="nEdgesOnCell"):
        swe.MPASMesh.from_netcdf(path)


def test_torch_restatement_matches_oracle_on_icosahedral_mesh(ico):
    """The differentiable restatement used as the adjoint's reference
    follows the oracle on pentagons too."""
    import torch
    import torch_swe as ts
    from helpers import ALL_TERMS, params_dict, rel_err
    L, ntr = 4, 2
    pd = params_dict(float(np.mean(ico["dc_edge"])), ntr)
    om = po.Mesh.from_arrays(ico, L)
    m = po.Model(om, **pd)
    state = meshgen.sphere_state(ico, L, ntr, 7)
    tend = m.new_state()
    m.compute_tendency(tend, state, 0.0)
    out = ts.tendency(ts.TorchMesh(om, L), pd, ALL_TERMS, *(torch.from_numpy(a) for a in state))
    for name, a, b in zip(("h", "vn", "tr"), out, tend):
        assert rel_err(a.numpy(), b) <= 1e-12, name
