"""Host tessellation: the reference's tessellation tests (src/space.rs:1453-1567) plus the identity of
the product's 1D mesh with the oracle's 1D mesh (the device 1D rebuild restates the same rule)."""
import numpy as np
import pytest

from hydro_rs_amd.ics import jittered_grid
from hydro_rs_amd.mesh import VoronoiMesh


def test_init_1d_known_answers():
    # src/space.rs:1453-1491
    loc = np.zeros((4, 3))
    loc[:, 0] = [0.25, 0.65, 0.75, 0.85]
    m = VoronoiMesh(1, [1.0, 1.0, 1.0], False, loc)
    np.testing.assert_allclose(m.volume, [0.45, 0.25, 0.1, 0.2], atol=1e-10)
    assert m.n_faces == 5
    np.testing.assert_array_equal(m.area, np.ones(5))
    for got, want in zip(m.centroid[:, 0], [0.0, 0.45, 0.7, 0.8, 1.0]):
        assert abs(got - want) <= 4 * np.spacing(max(abs(want), 1e-300)) or got == want


@pytest.mark.parametrize("periodic", [False, True])
def test_1d_matches_oracle(ora, periodic):
    rng = np.random.default_rng(0)
    n = 300
    x = (np.arange(n) + 0.5) / n + 0.01 / n * rng.uniform(-1, 1, n)
    loc = np.zeros((n, 3))
    loc[:, 0] = x
    m = VoronoiMesh(1, [1.0, 1.0, 1.0], periodic, loc)
    o = ora.voronoi_1d(x, 1.0, periodic)
    assert o["n_faces"] == m.n_faces
    for k in ("area", "normal", "centroid", "shift", "left", "right", "volume", "cell_centroid", "conn", "offset",
              "count"):
        np.testing.assert_array_equal(np.ravel(getattr(m, k)), o[k], err_msg=k)


@pytest.mark.parametrize("dim,n,periodic,box", [
    (2, 256, False, (2.0, 2.0, 1.0)),
    (2, 1024, True, (1.0, 1.0, 1.0)),
    (3, 512, False, (2.0, 2.0, 2.0)),
    (3, 1728, True, (1.0, 1.0, 1.0)),
])
def test_tessellation_invariants(dim, n, periodic, box):
    # src/space.rs:1493-1567: total volume = box volume, each cell's face area exceeds the isoperimetric
    # bound, and the generator lies within half the search radius of its centroid.
    rng = np.random.default_rng(dim * 100 + n)
    side = round(n ** (1.0 / dim))
    loc = jittered_grid(dim, side, box, 0.1, rng)
    m = VoronoiMesh(dim, box, periodic, loc)
    V = float(np.prod(box[:dim]))
    assert abs(m.volume.sum() - V) <= 1e-12 * V
    areas = np.zeros(m.n_cells)
    np.add.at(areas, m.left, m.area)
    inner = m.right >= 0
    np.add.at(areas, m.right[inner], m.area[inner])
    if dim == 2:
        iso = 2 * np.pi * np.sqrt(m.volume / np.pi)
    else:
        iso = 4 * np.pi * (3.0 / (4 * np.pi) * m.volume) ** (2.0 / 3.0)
    assert np.all(areas > iso)
    assert np.all(np.linalg.norm(loc - m.cell_centroid, axis=1) < 0.5 * m.search_radius)
    # structure: every face touches its cells' CSR, normals are unit, walls only without periodicity
    assert np.allclose(np.linalg.norm(m.normal, axis=1), 1.0)
    assert periodic == bool(np.all(m.right >= 0))
    assert m.conn.size == m.n_faces + int(inner.sum())
    for i in rng.integers(0, m.n_cells, 20):
        fs = m.conn[m.offset[i]:m.offset[i] + m.count[i]]
        assert np.all((m.left[fs] == i) | (m.right[fs] == i))
        assert np.all(np.diff(fs) > 0)


def test_linear_field_divergence_theorem():
    """Sum over a cell's faces of area * normal vanishes (closed cells), periodic 3D."""
    rng = np.random.default_rng(5)
    loc = jittered_grid(3, 10, (1.0, 1.0, 1.0), 0.1, rng)
    m = VoronoiMesh(3, (1.0, 1.0, 1.0), True, loc)
    s = np.zeros((m.n_cells, 3))
    np.add.at(s, m.left, m.area[:, None] * m.normal)
    np.add.at(s, m.right, -m.area[:, None] * m.normal)
    assert np.max(np.abs(s)) < 1e-12


@pytest.mark.parametrize("periodic", [True, False])
def test_3d_exact_lattice_known_answers(periodic):
    """Generators on an exact cubic lattice: every cell is the cube around its generator (volume w^3,
    centroid at the generator), with six faces of area w^2 and axis normals; 3 N faces on a periodic box
    (each shared), 3 N + 3 side^2 with walls."""
    side = 5
    w = 1.0 / side
    g = (np.arange(side) + 0.5) * w
    loc = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    m = VoronoiMesh(3, (1.0, 1.0, 1.0), periodic, loc)
    n = side ** 3
    assert m.n_faces == (3 * n if periodic else 3 * n + 3 * side * side)
    assert np.allclose(m.volume, w ** 3, rtol=1e-12, atol=0)
    assert np.allclose(m.cell_centroid, loc, rtol=0, atol=1e-14)
    assert np.allclose(m.area, w * w, rtol=1e-12, atol=0)
    assert np.allclose(np.abs(m.normal).max(axis=1), 1.0) and np.allclose(np.abs(m.normal).sum(axis=1), 1.0)
    assert np.all(m.count == 6)
    # each interior face sits halfway between its generators (through the periodic image where shifted)
    inner = m.right >= 0
    mid = 0.5 * (loc[m.left[inner]] + loc[m.right[inner]] + m.shift[inner])
    assert np.allclose(m.centroid[inner], mid, rtol=0, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("periodic", [True, False])
def test_partial_build_is_the_full_build_around_the_masked_cells(periodic, dim):
    """Voronoi::build_partial (space.rs:236-245): the masked cells' geometry equals the full build's, the
    faces are exactly the full build's faces with a masked cell (each from a masked cell's polygon: within
    1e-12 of the full build's, and bit for bit when both cells are masked), every cell's CSR lists exactly
    its faces, and unmasked cells are unbuilt."""
    rng = np.random.default_rng(11)
    box = np.array([1.0, 1.0, 1.0])
    loc = jittered_grid(dim, 20 if dim == 2 else 7, box, 0.3, rng)
    n = loc.shape[0]
    mask = rng.uniform(size=n) < 0.3
    full = VoronoiMesh(dim, box, periodic, loc)
    part = VoronoiMesh(dim, box, periodic, loc, mask=mask)
    for k in ("volume", "search_radius"):
        assert np.array_equal(getattr(part, k)[mask], getattr(full, k)[mask]), k
        assert np.all(getattr(part, k)[~mask] == 0.0), k
    assert np.array_equal(part.cell_centroid[mask], full.cell_centroid[mask])

    def key(m, f):
        a, b = int(m.left[f]), int(m.right[f])
        return (min(a, b), max(a, b), tuple(np.round(m.shift[f] * (1 if a <= b else -1), 12))) if b >= 0 else (a, -1, f"w{m.normal[f].round(6).tolist()}")

    fk = {key(full, f): f for f in range(full.n_faces)}
    touched = {k for k in fk if mask[k[0]] or (k[1] >= 0 and mask[k[1]])}
    pk = {key(part, f): f for f in range(part.n_faces)}
    assert set(pk) == touched
    for k, f in pk.items():
        g = fk[k]
        both = mask[k[0]] and k[1] >= 0 and mask[k[1]]
        assert abs(part.area[f] - full.area[g]) <= 1e-12 * full.area[g]
        if both:
            assert part.area[f] == full.area[g] and np.array_equal(part.centroid[f], full.centroid[g])
    for i in range(n):
        mine = sorted(part.conn[part.offset[i]: part.offset[i] + part.count[i]])
        want = sorted(f for f in range(part.n_faces) if part.left[f] == i or part.right[f] == i)
        assert mine == want
