"""Test fixtures for the general-mesh path: spherical icosahedral (geodesic)
Voronoi meshes in the reference's MPAS conventions, the TRiSK
edge-on-edge weights of any such mesh, and an MPAS NetCDF writer.

Conventions (the reference's planar mesh follows the same ones,
src/mesh/planar_hexagonal_mesh.cpp:169-292):

* edges_on_cell runs counterclockwise seen from outside the sphere, and edge
  j of a cell lies between its vertices j and j+1;
* the normal of edge e points from cells_on_edge[e, 0] to cells_on_edge[e, 1],
  the tangent k x n from vertices_on_edge[e, 0] to vertices_on_edge[e, 1];
* kiteareas_on_vertex[v, l] is the part of vertex v's triangle that lies in
  cell cells_on_vertex[v, l]; area_cell and area_triangle are sums of kites;
* edges_on_edge lists the other edges of cells_on_edge[e, 0] counterclockwise
  from e, then those of cells_on_edge[e, 1].

An icosahedron refined `level` times has 10*4**level + 2 cells: twelve
pentagons (five edges, nine edges-on-edge next to them), the rest hexagons.
"""
from __future__ import annotations

import numpy as np

MAXEDGES = 6


def trisk_weights(a: dict):
    """nedges_on_edge, edges_on_edge[ne, 12], weights_on_edge[ne, 12] of a
    Voronoi C-grid given its cell/edge/vertex connectivity, kite areas and
    edge lengths (0-based, reference names).

    Weight of edge e' of cell c in the tangential velocity at e
    (Thuburn, Ringler, Skamarock & Klemp 2009, J. Comput. Phys. 228:8321,
    eq. 33):  w = s(c,e) s(c,e') (1/2 - R) dv_e' / dc_e, where s(c,x) is +1
    when c is the first cell of x and R sums the kite-area fractions of c
    at the vertices passed going counterclockwise from e to e'."""
    noc, eoc, voc = a["nedges_on_cell"], a["edges_on_cell"], a["vertices_on_cell"]
    coe, cov, kites = a["cells_on_edge"], a["cells_on_vertex"], a["kiteareas_on_vertex"]
    area, dc, dv = a["area_cell"], a["dc_edge"], a["dv_edge"]
    ne = len(coe)
    neoe = np.zeros(ne, np.int32)
    eoe = np.zeros((ne, 2 * MAXEDGES), np.int32)
    w = np.zeros((ne, 2 * MAXEDGES))

    def kite(c, v):
        (l,) = np.nonzero(cov[v] == c)[0]
        return kites[v, l]

    def sign(c, e):
        return 1.0 if coe[e, 0] == c else -1.0

    for e in range(ne):
        n = 0
        for c in coe[e]:
            m = noc[c]
            j0 = list(eoc[c, :m]).index(e)
            r = 0.0
            for i in range(1, m):
                jj = (j0 + i) % m
                r += kite(c, voc[c, jj]) / area[c]
                ep = eoc[c, jj]
                eoe[e, n] = ep
                w[e, n] = sign(c, e) * sign(c, ep) * (0.5 - r) * dv[ep] / dc[e]
                n += 1
        neoe[e] = n
    return neoe, eoe, w


def _unit(x):
    return x / np.linalg.norm(x, axis=-1, keepdims=True)


def _angle(p, q):
    """Great-circle angle between unit vectors (rows)."""
    return np.arctan2(np.linalg.norm(np.cross(p, q), axis=-1), np.sum(p * q, axis=-1))


def _tri_area(a, b, c):
    """Signed area of the spherical triangle a, b, c on the unit sphere
    (positive counterclockwise seen from outside; Eriksson 1990)."""
    num = np.sum(a * np.cross(b, c), axis=-1)
    den = 1 + np.sum(a * b, axis=-1) + np.sum(b * c, axis=-1) + np.sum(c * a, axis=-1)
    return 2 * np.arctan2(num, den)


def _geodesic(level: int):
    t = (1 + 5 ** 0.5) / 2
    pts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
           (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    P = [np.array(p, float) / np.linalg.norm(p) for p in pts]
    for _ in range(level):
        mid = {}

        def m(i, j):
            k = (min(i, j), max(i, j))
            if k not in mid:
                mid[k] = len(P)
                P.append(_unit(P[i] + P[j]))
            return mid[k]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    P = np.array(P)
    T = np.array(faces, np.int64)
    # counterclockwise seen from outside
    n = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
    flip = np.sum(n * (P[T[:, 0]] + P[T[:, 1]] + P[T[:, 2]]), axis=1) < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    return P, T


def icosahedral_mesh(level: int = 2, radius: float = 6.371e6, seed: int = 0) -> dict:
    """MPAS arrays (reference names, 0-based) of the Voronoi dual of an
    icosahedron refined `level` times, on a sphere of the given radius.
    `seed` shuffles the cell, edge and vertex numbering and the first/second
    cell of each edge, so that edge signs of both kinds occur everywhere."""
    rng = np.random.default_rng(seed)
    P, T = _geodesic(level)
    # shuffle cell and vertex numbering
    cp = rng.permutation(len(P))
    P = P[np.argsort(cp)]
    T = cp[T]
    T = T[rng.permutation(len(T))]
    nc, nv = len(P), len(T)
    V = _unit(np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]]))

    # edges: triangle on the left of each directed cell pair
    left = {}
    for t, (a, b, c) in enumerate(T):
        for i, j in ((a, b), (b, c), (c, a)):
            left[(i, j)] = t
    pairs = sorted({(min(i, j), max(i, j)) for i, j in left})
    pairs = [pairs[i] for i in rng.permutation(len(pairs))]
    pairs = [(j, i) if rng.random() < 0.5 else (i, j) for i, j in pairs]
    ne = len(pairs)
    eid = {}
    coe = np.zeros((ne, 2), np.int32)
    voe = np.zeros((ne, 2), np.int32)
    for e, (i, j) in enumerate(pairs):
        eid[(i, j)] = eid[(j, i)] = e
        coe[e] = (i, j)
        # tangent k x n points to the left of i -> j
        voe[e] = (left[(j, i)], left[(i, j)])

    # cells: neighbours counterclockwise; vertex j sits between edges j-1, j
    succ = [dict() for _ in range(nc)]
    tri_of = {}
    for t, (a, b, c) in enumerate(T):
        for i, j, k in ((a, b, c), (b, c, a), (c, a, b)):
            succ[i][j] = k
            tri_of[(i, j, k)] = t
    noc = np.zeros(nc, np.int32)
    eoc = np.zeros((nc, MAXEDGES), np.int32)
    coc = np.zeros((nc, MAXEDGES), np.int32)
    voc = np.zeros((nc, MAXEDGES), np.int32)
    for i in range(nc):
        nb = [min(succ[i])]
        while succ[i][nb[-1]] != nb[0]:
            nb.append(succ[i][nb[-1]])
        m = len(nb)
        noc[i] = m
        for j in range(m):
            coc[i, j] = nb[j]
            eoc[i, j] = eid[(i, nb[j])]
            voc[i, j] = tri_of[(i, nb[j - 1], nb[j])]

    # vertices: cells in triangle order, edges (a,b), (b,c), (c,a)
    cov = T.astype(np.int32)
    eov = np.array([[eid[(a, b)], eid[(b, c)], eid[(c, a)]] for a, b, c in T], np.int32)
    kites = np.zeros((nv, 3))
    for l in range(3):
        c = P[T[:, l]]
        m_next = _unit(c + P[T[:, (l + 1) % 3]])
        m_prev = _unit(c + P[T[:, (l + 2) % 3]])
        kites[:, l] = (_tri_area(c, m_next, V) + _tri_area(c, V, m_prev)) * radius ** 2
    area_triangle = kites.sum(axis=1)
    area_cell = np.zeros(nc)
    for l in range(3):
        np.add.at(area_cell, T[:, l], kites[:, l])

    dc = _angle(P[coe[:, 0]], P[coe[:, 1]]) * radius
    dv = _angle(V[voe[:, 0]], V[voe[:, 1]]) * radius
    xe = _unit(P[coe[:, 0]] + P[coe[:, 1]])
    east = _unit(np.stack([-xe[:, 1], xe[:, 0], np.zeros(ne)], axis=1))
    north = np.cross(xe, east)
    nrm = P[coe[:, 1]] - P[coe[:, 0]]
    angle = np.arctan2(np.sum(nrm * north, axis=1), np.sum(nrm * east, axis=1))

    a = dict(nedges_on_cell=noc, edges_on_cell=eoc, cells_on_cell=coc, vertices_on_cell=voc,
             cells_on_edge=coe, vertices_on_edge=voe, edges_on_vertex=eov, cells_on_vertex=cov,
             kiteareas_on_vertex=kites, area_cell=area_cell, area_triangle=area_triangle,
             dc_edge=dc, dv_edge=dv, angle_edge=angle)
    a["nedges_on_edge"], a["edges_on_edge"], a["weights_on_edge"] = trisk_weights(a)
    for name, X in (("cell", P), ("edge", xe), ("vertex", V)):
        a["x_" + name], a["y_" + name], a["z_" + name] = (radius * X[:, i] for i in range(3))
        a["lat_" + name] = np.arcsin(np.clip(X[:, 2], -1, 1))
        a["lon_" + name] = np.arctan2(X[:, 1], X[:, 0])
    return a


# MPAS NetCDF name, storage kind and dimensions of the reference's mesh
# fields; "index" entries are stored 1-based with 0 meaning "none"
C, E, V = "nCells", "nEdges", "nVertices"
NETCDF_VARS = {
    "nedges_on_cell": ("nEdgesOnCell", "count", (C,)),
    "edges_on_cell": ("edgesOnCell", "index", (C, "maxEdges")),
    "cells_on_cell": ("cellsOnCell", "index", (C, "maxEdges")),
    "vertices_on_cell": ("verticesOnCell", "index", (C, "maxEdges")),
    "nedges_on_edge": ("nEdgesOnEdge", "count", (E,)),
    "edges_on_edge": ("edgesOnEdge", "index", (E, "maxEdges2")),
    "weights_on_edge": ("weightsOnEdge", "real", (E, "maxEdges2")),
    "cells_on_edge": ("cellsOnEdge", "index", (E, "TWO")),
    "vertices_on_edge": ("verticesOnEdge", "index", (E, "TWO")),
    "edges_on_vertex": ("edgesOnVertex", "index", (V, "vertexDegree")),
    "cells_on_vertex": ("cellsOnVertex", "index", (V, "vertexDegree")),
    "area_cell": ("areaCell", "real", (C,)),
    "dc_edge": ("dcEdge", "real", (E,)),
    "dv_edge": ("dvEdge", "real", (E,)),
    "angle_edge": ("angleEdge", "real", (E,)),
    "area_triangle": ("areaTriangle", "real", (V,)),
    "kiteareas_on_vertex": ("kiteAreasOnVertex", "real", (V, "vertexDegree")),
    "mesh_scaling_del2": ("meshScalingDel2", "real", (E,)),
    "mesh_scaling_del4": ("meshScalingDel4", "real", (E,)),
    "max_level_edge_top": ("maxLevelEdgeTop", "count", (E,)),
}
for _ent, _dim in (("cell", C), ("edge", E), ("vertex", V)):
    for _q, _nc in (("x", "x"), ("y", "y"), ("z", "z"), ("lat", "lat"), ("lon", "lon")):
        NETCDF_VARS[f"{_q}_{_ent}"] = (f"{_nc}{_ent.capitalize()}", "real", (_dim,))

# row counts of the padded index arrays
_COUNTS = {"edges_on_cell": "nedges_on_cell", "cells_on_cell": "nedges_on_cell",
           "vertices_on_cell": "nedges_on_cell", "edges_on_edge": "nedges_on_edge"}


def write_mpas_netcdf(path: str, a: dict, max_edges: int = MAXEDGES, nlayers: int | None = None,
                      sphere_radius: float | None = None) -> None:
    """Writes the arrays as an MPAS mesh file (NetCDF classic, 64-bit offsets),
    with 1-based indices and rows padded with 0 to maxEdges / maxEdges2."""
    from scipy.io import netcdf_file

    sizes = {C: len(a["nedges_on_cell"]), E: len(a["nedges_on_edge"]),
             V: len(a["edges_on_vertex"]), "maxEdges": max_edges, "maxEdges2": 2 * max_edges,
             "TWO": 2, "vertexDegree": 3}
    with netcdf_file(path, "w", version=2) as f:
        f.on_a_sphere = b"YES" if sphere_radius else b"NO"
        f.sphere_radius = float(sphere_radius or 0.0)
        for d, n in sizes.items():
            f.createDimension(d, n)
        if nlayers is not None:
            f.createDimension("nVertLevels", nlayers)
        for name, x in a.items():
            if name not in NETCDF_VARS or x is None:
                continue
            ncname, kind, dims = NETCDF_VARS[name]
            x = np.asarray(x)
            y = np.zeros(tuple(sizes[d] for d in dims), np.float64 if kind == "real" else np.int32)
            if kind == "index":
                x = x + 1
                if name in _COUNTS:
                    cnt = np.asarray(a[_COUNTS[name]])
                    x = np.where(np.arange(x.shape[1])[None, :] < cnt[:, None], x, 0)
            if x.ndim == 1:
                y[:] = x
            else:
                y[:, :x.shape[1]] = x
            v = f.createVariable(ncname, "d" if kind == "real" else "i", dims)
            v[:] = y


def sphere_state(a: dict, nlayers: int, ntracers: int, seed: int = 0):
    """Smooth layer-dependent state with noise on a mesh from
    icosahedral_mesh: h ~ 1000, vn ~ 1, thickness-weighted tracers."""
    rng = np.random.default_rng(seed)
    lat, lon = np.asarray(a["lat_cell"]), np.asarray(a["lon_cell"])
    k = np.arange(nlayers)[None, :]
    h = 1000 + 20 * (np.sin(2 * lon) * np.cos(lat))[:, None] * (1 + 0.01 * k)
    h = h + 5 * rng.standard_normal(h.shape)
    late, lone = np.asarray(a["lat_edge"]), np.asarray(a["lon_edge"])
    ang = np.asarray(a["angle_edge"])
    u = (np.cos(late) * np.cos(ang) + 0.3 * np.sin(3 * lone) * np.sin(ang))[:, None]
    vn = u * (1 - 0.005 * k) + 0.3 * rng.standard_normal((len(ang), nlayers))
    tr = np.empty((ntracers, len(lat), nlayers))
    for l in range(ntracers):
        base = 1.5 + np.sin((l + 1) * lon) * np.cos(lat)
        tr[l] = h * (base[:, None] * (1 + 0.02 * k) + 0.01 * rng.standard_normal(h.shape))
    return np.ascontiguousarray(h), np.ascontiguousarray(vn), np.ascontiguousarray(tr)
