"""Mesh models on the host and in the oracle (FinFET, 4080 triangles).

- Facets, cell grid and grid lookups of the host equal the reference's
  SimulationDomain built from the same STL, bit for bit
  (tests/golden/domain_finfet.npz, written by make_golden.py --mesh).
- The oracle's restatement of SurfaceMesh::trace_boundary_intersection,
  contains_point and nearest_facet equals the reference on 1500 interior rays,
  1100 rays leaving surface points and 4800 points on, near and away from the
  surface (points, times and facets bit for bit).
- Reservoir draws on the mesh's bottom facet (mode, time and area-weighted
  triangle point) equal the reference's.
- Independently of the reference: the procedural FinFET is closed and
  manifold with the closed-form volume; the bounding-volume hierarchies hold
  every face once and contain it; the star tetrahedra tile the volume; the
  oracle's volume sampling is uniform (cell counts against cell volumes).
"""
import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt


@pytest.fixture(scope="module")
def finfet():
    return rt.Problem(rt.config_text("finfet.toml"), input_directory=rt.CONFIGS, mode_sampling=1)


@pytest.fixture(scope="module")
def mesh(finfet):
    _, _, g, _, _, _ = finfet.views()
    assert g.kind == 1
    return g.mesh.contents


def test_mesh_domain_bit_exact(finfet, golden):
    g = golden("domain_finfet.npz")
    f = finfet.facets()
    for name in ["normals", "centroids", "areas", "shifts"]:
        assert np.array_equal(f[name], g[name]), name
    assert f["conditions"] == bytes(g["conditions"]).decode() == "PPTRPRRRRRP"
    assert np.array_equal(f["faces"], g["faces"])
    centers, volumes = finfet.cells()
    assert np.array_equal(centers, g["centers"]) and np.array_equal(volumes, g["volumes"])
    assert len(volumes) == 1080
    assert finfet.scalars()["volume"] == float(g["volume"])
    cells = np.array([O.grid_index(finfet, x) for x in g["grid_pos"]])
    assert np.array_equal(cells, g["grid_cell"])


def test_mesh_trace_bit_exact(finfet, golden):
    g = golden("domain_finfet.npz")
    pts, t, facet = O.trace_mesh(finfet, g["mesh_trace_pos"], g["mesh_trace_vel"])
    assert np.array_equal(facet, g["mesh_trace_facet"])
    assert np.array_equal(t, g["mesh_trace_t"])
    assert np.array_equal(pts, g["mesh_trace_point"])
    assert (facet >= 0).all()  # closed surface: every interior ray hits it
    # The reference's box-geometry rays of domain_vectors go through the same trace.
    pts, t, facet = O.trace_mesh(finfet, g["trace_pos"], g["trace_vel"])
    assert np.array_equal(facet, g["trace_facet"])
    assert np.array_equal(t[facet >= 0], g["trace_t"][facet >= 0])


def test_mesh_containment_and_nearest_facet_bit_exact(finfet, golden):
    g = golden("domain_finfet.npz")
    inside = O.contains_mesh(finfet, g["mesh_contains_pos"])
    assert np.array_equal(inside, g["mesh_contains"].astype(bool))
    assert 0.2 < inside.mean() < 0.9  # both answers are exercised
    assert np.array_equal(O.nearest_facet(finfet, g["mesh_nearest_pos"]), g["mesh_nearest"])


def test_mesh_reservoir_draws_bit_exact(finfet, golden):
    g = golden("reservoir_finfet.npz")
    s, m, geo, i, rough, res = finfet.views()
    assert len(res) == 1
    assert res[0].initial_count == int(g["r0_initial"])
    u = g["r0_u"]
    modes = [O.reservoir_mode(finfet, 0, u[k, 0], 1) for k in range(len(u))]
    assert np.array_equal(modes, g["r0_mode"])
    su = g["r0_surf_u"]
    pts = np.array([O.surface_point(finfet, 0, *su[k]) for k in range(len(su))])
    assert np.array_equal(pts, g["r0_surf_point"])


def test_finfet_is_closed_manifold_with_closed_form_volume(mesh, finfet):
    faces = rt.arr(mesh.faces, 3 * mesh.n_faces, np.int32).reshape(-1, 3)
    assert mesh.n_faces == 4080
    edges = {}
    for tri in faces:
        for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            edges[(a, b)] = edges.get((a, b), 0) + 1
    # Each directed edge once and its reverse once: closed, manifold, consistently wound.
    assert all(c == 1 for c in edges.values())
    assert all((b, a) in edges for (a, b) in edges)
    # 200 x 100 x 50 nm substrate + 10 x 40 x 100 nm fin, in A^3.
    assert finfet.scalars()["volume"] == pytest.approx((200 * 100 * 50 + 10 * 40 * 100) * 1e3, rel=1e-12)
    v = rt.arr(mesh.vertices, 3 * mesh.n_vertices).reshape(-1, 3)
    tri = v[faces]
    signed = np.einsum("ij,ij->i", tri[:, 0], np.cross(tri[:, 1], tri[:, 2])).sum() / 6.0
    assert signed == pytest.approx(finfet.scalars()["volume"], rel=1e-12)  # outward winding


def walk(nodes, n_nodes):
    """Leaves reached from the root, each with its chain of boxes."""
    out, stack = [], [(0, [])]
    while stack:
        i, chain = stack.pop()
        node = nodes[i]
        box = (np.array(node.bmin[:]), np.array(node.bmax[:]))
        if node.count > 0:
            out.append((node.index, node.count, chain + [box]))
        else:
            assert i + 1 < n_nodes and i < node.index < n_nodes
            stack.append((i + 1, chain + [box]))
            stack.append((node.index, chain + [box]))
    return out


def test_bvh_holds_each_item_once_inside_its_boxes(mesh):
    order = rt.arr(mesh.face_order, mesh.n_faces, np.int32)
    assert np.array_equal(np.sort(order), np.arange(mesh.n_faces))
    bounds = rt.arr(mesh.face_bounds, 6 * mesh.n_faces).reshape(-1, 6)
    seen = np.zeros(mesh.n_faces, int)
    depth = 0
    for first, count, chain in walk(mesh.face_nodes, mesh.n_face_nodes):
        assert count <= 4
        depth = max(depth, len(chain))
        for j in range(first, first + count):
            f = order[j]
            seen[f] += 1
            for lo, hi in chain:
                assert (lo <= bounds[f, :3]).all() and (bounds[f, 3:] <= hi).all()
    assert (seen == 1).all()
    assert depth <= 48
    sorder = rt.arr(mesh.simplex_order, mesh.n_simplices, np.int32)
    assert np.array_equal(np.sort(sorder), np.arange(mesh.n_simplices))


def test_star_tetrahedra_tile_the_volume(mesh, finfet):
    assert mesh.n_simplices > 0  # the FinFET is star-shaped about its vertex mean
    vol = rt.arr(mesh.simplex_volume, mesh.n_simplices)
    assert vol.sum() == pytest.approx(finfet.scalars()["volume"], rel=1e-9)
    sv = rt.arr(mesh.simplex_vertices, 12 * mesh.n_simplices).reshape(-1, 4, 3)
    e = sv[:, 1:] - sv[:, :1]
    assert np.allclose(np.abs(np.linalg.det(e)) / 6.0, vol, rtol=1e-12)
    inv = rt.arr(mesh.simplex_inverse, 9 * mesh.n_simplices).reshape(-1, 3, 3)
    assert np.allclose(inv @ np.transpose(e, (0, 2, 1)), np.eye(3), atol=1e-9)


def test_volume_sampling_is_uniform(finfet):
    """Oracle volume draws (Philox, star tetrahedra + Dirichlet weights): every
    point is inside; the fin holds its volume fraction (40 000 of 1 040 000
    nm^3); over the 1000 substrate cells (10 nm cubes wholly inside) the counts
    are uniform (chi-square)."""
    ids = np.arange(40000)
    pts, um = O.sample_volume(finfet, 12345, 0, ids)
    assert O.contains_mesh(finfet, pts[:4000]).all()
    assert ((um >= 0) & (um < 1)).all()
    fin = pts[:, 2] > 500.0
    frac = 40000.0 / 1040000.0
    assert abs(fin.mean() - frac) < 5 * np.sqrt(frac * (1 - frac) / len(pts))
    sub = pts[~fin]
    cells = np.array([O.grid_index(finfet, x) for x in sub])
    centers, _ = finfet.cells()
    lower = np.flatnonzero(centers[:, 2] < 500.0)
    assert len(lower) == 1000
    counts = np.bincount(cells, minlength=len(centers))[lower]
    assert counts.sum() == len(sub)
    expect = len(sub) / 1000.0
    chi2 = ((counts - expect) ** 2 / expect).sum()
    assert abs(chi2 - 999) < 5 * np.sqrt(2 * 999), chi2


def point_triangle_distance(p, a, b, c):
    """Distance from point p to each triangle (a, b, c rows): Voronoi regions
    of the triangle (Ericson 5.1.5), vectorized over triangles."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = (ab * ap).sum(1), (ac * ap).sum(1)
    bp, cp = p - b, p - c
    d3, d4 = (ab * bp).sum(1), (ac * bp).sum(1)
    d5, d6 = (ab * cp).sum(1), (ac * cp).sum(1)
    vc, vb, va = d1 * d4 - d3 * d2, d5 * d2 - d1 * d6, d3 * d6 - d5 * d4
    with np.errstate(divide="ignore", invalid="ignore"):
        den = 1.0 / (va + vb + vc)
        v, w = vb * den, vc * den
        cases = [
            (d1 <= 0) & (d2 <= 0), (d3 >= 0) & (d4 <= d3), (vc <= 0) & (d1 >= 0) & (d3 <= 0),
            (d6 >= 0) & (d5 <= d6), (vb <= 0) & (d2 >= 0) & (d6 <= 0),
            (va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0)]
        vals = [(0, 0), (1, 0), (d1 / (d1 - d3), 0), (0, 1), (0, d2 / (d2 - d6)),
                (1 - (d4 - d3) / ((d4 - d3) + (d5 - d6)), (d4 - d3) / ((d4 - d3) + (d5 - d6)))]
    done = np.zeros(len(a), bool)
    for cond, (vv, ww) in zip(cases, vals):
        pick = cond & ~done
        v = np.where(pick, vv, v)
        w = np.where(pick, ww, w)
        done |= cond
    q = a + v[:, None] * ab + w[:, None] * ac
    return np.linalg.norm(p - q, axis=1)


def test_distance_field_bounds_the_surface_distance(mesh, finfet):
    """Every voxel's stored bound is at most the brute-force distance from points
    of that voxel to the nearest face, and the bound is not vacuous (it lets
    most interior points skip the face search for an 18 A flight)."""
    dims, h, unit = list(mesh.field_dims), mesh.field_h, mesh.field_unit
    assert np.prod(dims) <= 4 << 20 and unit == 0.25
    field = np.ctypeslib.as_array(mesh.field, shape=(int(np.prod(dims)),))
    _, _, g, _, _, _ = finfet.views()
    lo, hi = np.array(g.bmin[:]), np.array(g.bmax[:])
    faces = rt.arr(mesh.faces, 3 * mesh.n_faces, np.int32).reshape(-1, 3)
    v = rt.arr(mesh.vertices, 3 * mesh.n_vertices).reshape(-1, 3)
    a, b, c = v[faces[:, 0]], v[faces[:, 1]], v[faces[:, 2]]
    rng = np.random.default_rng(3)
    pts = lo + (hi - lo) * rng.random((1500, 3))
    # points within a few angstroms of faces as well
    f = rng.integers(0, len(faces), 500)
    r = rng.random((500, 2))
    s = np.sqrt(r[:, :1])
    near = (1 - s) * a[f] + s * (1 - r[:, 1:]) * b[f] + s * r[:, 1:] * c[f] + rng.normal(size=(500, 3)) * 3.0
    pts = np.clip(np.concatenate([pts, near]), lo, hi)
    idx = np.minimum(((pts - lo) / h).astype(int), np.array(dims) - 1)
    bound = field[(idx[:, 0] * dims[1] + idx[:, 1]) * dims[2] + idx[:, 2]] * unit
    dist = np.array([point_triangle_distance(p, a, b, c).min() for p in pts])
    assert (bound <= dist + 1e-9).all(), np.max(bound - dist)
    inside = O.contains_mesh(finfet, pts[:1500])
    assert np.mean(bound[:1500][inside] > 18.0 + 0.01) > 0.6


def test_distance_field_sides_match_containment(mesh, finfet):
    """Voxels clear of the surface carry the side of the surface they lie on:
    the flag equals the oracle's contains_point at random points of them."""
    dims, h = list(mesh.field_dims), mesh.field_h
    n = int(np.prod(dims))
    bound = np.ctypeslib.as_array(mesh.field, shape=(n,))
    inside = np.ctypeslib.as_array(mesh.field_inside, shape=(n,))
    _, _, g, _, _, _ = finfet.views()
    lo = np.array(g.bmin[:])
    rng = np.random.default_rng(5)
    clear = np.flatnonzero(bound > 0)
    pick = rng.choice(clear, 3000, replace=False)
    ijk = np.stack([pick // (dims[1] * dims[2]), (pick // dims[2]) % dims[1], pick % dims[2]], axis=1)
    pts = lo + (ijk + rng.random((len(pick), 3))) * h
    want = O.contains_mesh(finfet, pts)
    assert np.array_equal(inside[pick].astype(bool), want)
    assert 0.05 < want.mean() < 0.95  # both sides are present (the bounding box holds outside space)
