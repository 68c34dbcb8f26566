"""DMPlex / Hdf5Serializer ingestion (abl_fv_mesh_dmplex, ablate_amd.hdf5) on the reference's own HDF5 files
(tests/golden/*.hdf5, copied data):

- 2D_Circle.hdf5 is the Hdf5Serializer output of `!ablate::domain::BoxMesh` faces [80, 80] on [-2, 2]^2
  (tests/integrationTests/inputs/levelSet/2D_Circle.yaml) with the volume fraction of the unit circle: its mesh must
  equal the native box mesh of the same shape cell for cell, and the saved field must integrate to about pi;
- 2D_Circle_Tri.hdf5 is the same box split into triangles (`simplex: true`);
- initializer.{2D,3D}.hdf5 carry the fields of the reference's Hdf5Initializer test, whose table
  (tests/golden/hdf5_initializer.json, from hdf5InitializerTests.cpp) gives the field values at points: the cell
  holding each point (the first in point order, as DMLocatePoints returns it) must store exactly those values.
Geometry is also checked against what it fixes: total volume and first moments, closure of every cell, boundary
normals summing to the area vectors of the domain's sides, ghost cells mirroring their interior cell."""
import json
from pathlib import Path

import numpy as np
import pytest

from ablate_amd.hdf5 import H5File
from ablate_amd.mesh import BoxMesh, DMPlexMesh, Hdf5Mesh

GOLDEN = Path(__file__).resolve().parent / "golden"
CIRCLE = GOLDEN / "2D_Circle.hdf5"
CIRCLE_TRI = GOLDEN / "2D_Circle_Tri.hdf5"
INIT2 = GOLDEN / "initializer.2D.hdf5"
INIT3 = GOLDEN / "initializer.3D.hdf5"
ORDERINGS = ["natural", "rcm", "morton"]


def closure(a):
    S = np.zeros((a.n_cells, 3))
    np.add.at(S, a.face_cells[:, 0], a.face_normal)
    np.add.at(S, a.face_cells[:, 1], -a.face_normal)
    return np.abs(S[: a.n_interior]).max()


def ghost_mirror_error(a):
    """max |x_ghost - reflection of x_interior across its face plane| over the boundary faces."""
    b = np.nonzero(a.face_bc >= 0)[0]
    L, G = a.face_cells[b, 0], a.face_cells[b, 1]
    n = a.face_normal[b]
    nn = (n * n).sum(axis=1, keepdims=True)
    xi, xf = a.cell_centroid[L], a.face_centroid[b]
    refl = xi + 2.0 * ((xf - xi) * n).sum(axis=1, keepdims=True) / nn * n
    return np.abs(refl - a.cell_centroid[G]).max(), np.abs(a.cell_volume[G] - a.cell_volume[L]).max()


def test_reader_reads_the_dmplex_datasets():
    h = H5File(CIRCLE)
    assert set(h.keys("/")) >= {"topology", "geometry", "labels", "fields"}
    cones = h.read("/topology/cones").ravel()
    # 6400 quadrilaterals (4 edges), 12960 edges (2 vertices), 6561 vertices
    assert np.bincount(cones).tolist() == [6561, 0, 12960, 0, 6400]
    assert h.read("/topology/cells").size == 4 * 6400 + 2 * 12960
    X = h.read("/geometry/vertices")
    assert X.shape == (6561, 2) and X.min() == -2.0 and X.max() == 2.0
    assert h.keys("/labels/Face Sets") == ["1", "2", "3", "4"]
    assert h.read("/fields/solution").shape == (2, 6400)
    with pytest.raises(KeyError):
        h.read("/no/such/dataset")


@pytest.mark.parametrize("ordering", ORDERINGS)
def test_circle_file_is_the_box_mesh(ordering):
    """Cell for cell (matched through the source numbering: DMPlex cell point = lexicographic box index) the same
    centroids and volumes as abl_fv_mesh_box; the same faces; sides labelled 1 (y = -2), 2 (x = 2), 3 (y = 2),
    4 (x = -2) as DMPlexCreateBoxMesh labels them."""
    m = Hdf5Mesh(CIRCLE, ordering)
    a = m.arrays
    box = BoxMesh([80, 80], [-2, -2], [2, 2], periodic=[False, False], ordering="natural")
    b = box.arrays
    assert (a.n_interior, a.n_cells, a.n_faces) == (b.n_interior, b.n_cells, b.n_faces) == (6400, 6720, 12960)
    src = m.source_cells()
    assert sorted(src.tolist()) == list(range(6400))
    assert np.abs(a.cell_centroid[: a.n_interior] - b.cell_centroid[src]).max() < 1e-14
    assert np.abs(a.cell_volume[: a.n_interior] - b.cell_volume[src]).max() < 1e-16
    key = lambda arr: np.round(np.concatenate([arr.face_centroid, np.abs(arr.face_normal)], axis=1), 12)
    ka, kb = key(a), key(b)
    assert np.array_equal(ka[np.lexsort(ka.T[::-1])], kb[np.lexsort(kb.T[::-1])])
    assert closure(a) < 1e-15
    side = {1: [0, -4, 0], 2: [4, 0, 0], 3: [0, 4, 0], 4: [-4, 0, 0]}
    for v, s in side.items():
        assert a.face_normal[a.face_bc == v].sum(axis=0) == pytest.approx(s, abs=1e-13)
    assert max(ghost_mirror_error(a)) < 1e-15


@pytest.mark.parametrize("ordering", ORDERINGS)
def test_triangle_file(ordering):
    m = Hdf5Mesh(CIRCLE_TRI, ordering)
    a = m.arrays
    NI = a.n_interior
    assert NI == 12800 and a.n_cells - NI == 320 and m.labels == {1: 1}
    assert np.allclose(a.cell_volume[:NI], 0.00125, rtol=1e-12)
    assert a.cell_volume[:NI].sum() == pytest.approx(16.0, rel=1e-13)
    assert np.abs((a.cell_volume[:NI, None] * a.cell_centroid[:NI]).sum(axis=0)).max() < 1e-12
    assert closure(a) < 1e-15
    bnd = a.face_bc == 1
    xf, n = a.face_centroid[bnd], a.face_normal[bnd]
    for axis, val, s in ((1, -2, [0, -4, 0]), (0, 2, [4, 0, 0]), (1, 2, [0, 4, 0]), (0, -2, [-4, 0, 0])):
        on = np.abs(xf[:, axis] - val) < 1e-12
        assert on.sum() == 80 and n[on].sum(axis=0) == pytest.approx(s, abs=1e-13)
    assert max(ghost_mirror_error(a)) < 1e-14


@pytest.mark.parametrize("path", [CIRCLE, CIRCLE_TRI], ids=["quad", "tri"])
def test_circle_volume_fraction(path):
    """The saved volume fraction of the unit circle: 1 in cells inside it, 0 in cells outside, pi in total."""
    m = Hdf5Mesh(path, "morton")
    a = m.arrays
    NI = a.n_interior
    vof = m.fields()
    assert vof.shape == (1, NI)
    vof = vof[0]
    assert np.all((vof >= 0) & (vof <= 1))
    assert (a.cell_volume[:NI] * vof).sum() == pytest.approx(np.pi, abs=2e-3)
    r = np.hypot(a.cell_centroid[:NI, 0], a.cell_centroid[:NI, 1])
    h = 4.0 / 80 * np.sqrt(2.0)  # largest distance from a centroid to a vertex of its cell
    assert np.all(vof[r < 1 - h] == 1.0) and np.all(vof[r > 1 + h] == 0.0)
    cut = np.abs(r - 1) < 0.01  # the circle passes through the disk of radius 0.0118 about the centroid: inside
    assert np.all((vof[cut] > 0) & (vof[cut] < 1))


def cell_boxes(a):
    """Axis-aligned extent of every interior cell of a box mesh, from its face centroids."""
    NI = a.n_interior
    lo = np.full((NI, 3), np.inf)
    hi = np.full((NI, 3), -np.inf)
    for s in range(2):
        c = a.face_cells[:, s]
        ok = c < NI
        np.minimum.at(lo, c[ok], a.face_centroid[ok])
        np.maximum.at(hi, c[ok], a.face_centroid[ok])
    return lo, hi


@pytest.mark.parametrize("ordering", ORDERINGS)
@pytest.mark.parametrize("fname", ["initializer.2D.hdf5", "initializer.3D.hdf5"])
def test_initializer_known_answers(fname, ordering):
    table = json.loads((GOLDEN / "hdf5_initializer.json").read_text())["files"][fname]
    m = Hdf5Mesh(GOLDEN / fname, ordering)
    a = m.arrays
    dim = a.dim
    lo, hi = cell_boxes(a)
    src = m.source_cells()
    for field, points in table.items():
        vals = m.fields(f"solution_{field}", group="cell_fields")
        for tp in points:
            x = np.array(tp["point"][:dim])
            inside = np.all((lo[:, :dim] - 1e-12 <= x) & (x <= hi[:, :dim] + 1e-12), axis=1)
            cands = np.nonzero(inside)[0]
            assert cands.size >= 1
            c = cands[np.argmin(src[cands])]  # the first cell in point order holding the point
            want = np.array(tp["expected"])
            # ASSERT_DOUBLE_EQ of the reference test: equal within 4 units in the last place
            assert np.all(np.abs(vals[:, c] - want) <= 4 * np.spacing(np.abs(want))), (field, tp, vals[:, c])


@pytest.mark.parametrize("fname,dim,vol", [("initializer.2D.hdf5", 2, 1.0), ("initializer.3D.hdf5", 3, 1.0)])
def test_initializer_meshes_geometry(fname, dim, vol):
    m = Hdf5Mesh(GOLDEN / fname)
    a = m.arrays
    NI = a.n_interior
    assert a.dim == dim and a.cell_volume[:NI].sum() == pytest.approx(vol, rel=1e-13)
    assert closure(a) < 1e-15
    assert max(ghost_mirror_error(a)) < 1e-15
    # the first stored field component is 10 + x of the cell centroid (both files), the second 100 + y
    f = m.fields()
    assert np.abs(f[0] - 10 - a.cell_centroid[:NI, 0]).max() < 1e-12
    assert np.abs(f[1] - 100 - a.cell_centroid[:NI, 1]).max() < 1e-12


def two_triangles():
    """A DMPlex of two triangles on the unit square: cells 0-1, vertices 2-5, edges 6-10."""
    cones = [[6, 7, 8], [8, 9, 10], [], [], [], [], [2, 3], [3, 5], [5, 2], [5, 4], [4, 2]]
    X = [[0, 0], [1, 0], [0, 1], [1, 1]]  # points 2, 3, 4, 5
    return cones, X


def test_hand_built_plex():
    cones, X = two_triangles()
    m = DMPlexMesh(2, cones, X, {7: [6], 8: [7], 9: [9], 10: [10]})
    a = m.arrays
    assert (a.n_interior, a.n_cells, a.n_faces) == (2, 6, 5)
    assert m.source_cells().tolist() == [0, 1]
    assert a.cell_volume[:2].tolist() == [0.5, 0.5]
    assert a.cell_centroid[0] == pytest.approx([2 / 3, 1 / 3, 0])
    assert a.cell_centroid[1] == pytest.approx([1 / 3, 2 / 3, 0])
    inner = np.nonzero(a.face_bc < 0)[0]
    assert inner.size == 1 and a.face_cells[inner[0]].tolist() == [0, 1]
    assert a.face_normal[inner[0]] == pytest.approx([-1, 1, 0])  # from cell 0 towards cell 1, |n| = sqrt(2)
    assert sorted(a.face_bc[a.face_bc >= 0].tolist()) == [7, 8, 9, 10]
    assert closure(a) < 1e-16


def test_plex_errors():
    cones, X = two_triangles()
    with pytest.raises(RuntimeError, match="Face Sets"):
        DMPlexMesh(2, cones, X, {7: [6]})  # unlabelled boundary faces
    with pytest.raises(RuntimeError, match="depth|interpolated|deeper"):
        DMPlexMesh(2, [[1, 2, 3], [], [], []], [[0, 0], [1, 0], [0, 1]])  # uninterpolated: a cell cone of vertices
    with pytest.raises(RuntimeError, match="out of range"):
        DMPlexMesh(2, [[6, 7, 99]] + cones[1:], X, {7: [6]})
    with pytest.raises(RuntimeError, match="loop|close"):
        bad = [list(c) for c in cones]
        bad[8] = [2, 4]  # the diagonal edge no longer joins the triangles' vertices consistently
        DMPlexMesh(2, bad, X, {7: [6], 8: [7], 9: [9], 10: [10]})


def test_plex_ghost_cells_are_rebuilt():
    """Ghost cells of DMPlexConstructGhostCells (points of depth dim whose cone is a single boundary face) are dropped
    and made again: the mesh equals the one built from the same plex without them."""
    cones, X = two_triangles()
    labels = {7: [6], 8: [7], 9: [9], 10: [10]}
    base = DMPlexMesh(2, cones, X, labels).arrays
    # append one ghost cell per boundary edge (points 11-14), cones of one face each
    with_ghosts = cones + [[6], [7], [9], [10]]
    g = DMPlexMesh(2, with_ghosts, X, labels).arrays
    assert (g.n_interior, g.n_cells, g.n_faces) == (base.n_interior, base.n_cells, base.n_faces)
    for name in ("face_cells", "face_normal", "face_centroid", "cell_centroid", "cell_volume", "face_bc"):
        assert np.array_equal(getattr(g, name), getattr(base, name)), name


def test_stored_point_order_is_undone(monkeypatch):
    """/topology/order gives the global number of each stored point: a file that stores the points of the
    two-triangle plex in a shuffled order yields the same mesh as the plex itself."""
    import ablate_amd.hdf5 as h5mod

    cones, X = two_triangles()
    perm = np.random.default_rng(3).permutation(len(cones))  # stored position p holds global point perm[p]
    data = {
        "/topology/order": perm.astype(np.int32)[:, None],
        "/topology/cones": np.array([len(cones[g]) for g in perm], dtype=np.int32)[:, None],
        "/topology/cells": np.array([q for g in perm for q in cones[g]], dtype=np.int32)[:, None],
        "/geometry/vertices": np.array(X, dtype=np.float64),
        "/labels/Face Sets/7/indices": np.array([[6]], dtype=np.int32),
        "/labels/Face Sets/8/indices": np.array([[7]], dtype=np.int32),
        "/labels/Face Sets/9/indices": np.array([[9]], dtype=np.int32),
        "/labels/Face Sets/10/indices": np.array([[10]], dtype=np.int32),
    }

    class FakeFile:
        def __init__(self, path):
            self.path = str(path)

        def read(self, path):
            return data[path]

        def has(self, path):
            return any(k.startswith(path) for k in data)

        def keys(self, path):
            pre = path.rstrip("/") + "/"
            return sorted({k[len(pre):].split("/")[0] for k in data if k.startswith(pre)})

    monkeypatch.setattr(h5mod, "H5File", FakeFile)
    a = Hdf5Mesh("shuffled.h5").arrays
    b = DMPlexMesh(2, cones, X, {7: [6], 8: [7], 9: [9], 10: [10]}).arrays
    for name in ("face_cells", "face_normal", "face_centroid", "cell_centroid", "cell_volume", "face_bc"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
