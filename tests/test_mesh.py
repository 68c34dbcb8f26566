"""Generated meshes follow the DMPlex FV geometry conventions: closed cells, oriented normals, reflected ghosts,
periodic images, consistent volumes."""
import numpy as np
import pytest

from ablate_amd.mesh import BoxMesh, TetBoxMesh


def _closure(m):
    a = m.arrays
    acc = np.zeros((a.n_cells, 3))
    np.add.at(acc, a.face_cells[:, 0], a.face_normal)
    np.add.at(acc, a.face_cells[:, 1], -a.face_normal)
    return acc[: a.n_interior]


@pytest.mark.parametrize("mesh", [
    BoxMesh([5, 4], [0, 0], [2, 1], [True, True], "natural"),
    BoxMesh([6, 5], [0, 0], [1, 1], [False, True], "brick"),
    BoxMesh([4, 5, 6], [0, 0, 0], [1, 2, 3], [True, False, True], "rcm"),
    BoxMesh([8, 8, 8], [0, 0, 0], [1, 1, 1], [True, True, True], "brick"),
    TetBoxMesh(3, 0.2, ordering="rcm"),
    TetBoxMesh(4, 0.0, ordering="natural"),
], ids=lambda m: type(m).__name__)
def test_geometry_conventions(mesh):
    a = mesh.arrays
    NI = a.n_interior
    # every interior cell is closed: sum of outward area normals = 0
    scale = np.abs(a.face_normal).max()
    assert np.abs(_closure(mesh)).max() <= 1e-12 * scale
    # volumes fill the box
    assert a.cell_volume[:NI].sum() == pytest.approx(
        np.prod(np.asarray(getattr(mesh, "upper", [1, 1, 1]))[: a.dim] - np.asarray(getattr(mesh, "lower", [0, 0, 0]))[: a.dim]),
        rel=1e-12)
    # normals point from the left cell to the right cell (periodic image via face_shift)
    shift = a.face_shift if a.face_shift is not None else np.zeros_like(a.face_normal)
    xl = a.cell_centroid[a.face_cells[:, 0]]
    xr = a.cell_centroid[a.face_cells[:, 1]] + shift
    assert np.all(np.einsum("fd,fd->f", xr - xl, a.face_normal) > 0)
    # left cell interior; right cells in range; each ghost has exactly one face
    assert np.all(a.face_cells[:, 0] < NI)
    cnt = np.bincount(a.face_cells[:, 1], minlength=a.n_cells)
    assert np.all(cnt[NI:] == 1)
    # ghost centroids are reflections of the interior centroid across the face plane, same volume
    b = np.where(a.face_cells[:, 1] >= NI)[0]
    if b.size:
        g = a.face_cells[b, 1]
        i = a.face_cells[b, 0]
        mid = 0.5 * (a.cell_centroid[g] + a.cell_centroid[i])
        nrm = a.face_normal[b] / np.linalg.norm(a.face_normal[b], axis=1, keepdims=True)
        off = np.einsum("fd,fd->f", mid - a.face_centroid[b], nrm)
        assert np.abs(off).max() <= 1e-12
        np.testing.assert_array_equal(a.cell_volume[g], a.cell_volume[i])
        assert a.face_bc is not None and np.all(a.face_bc[b] >= 0)
        assert np.all(a.face_bc[a.face_cells[:, 1] < NI] == -1)
    # faces are numbered by their left cell
    assert np.all(np.diff(a.face_cells[:, 0]) >= 0)


def test_box_counts_and_orderings_agree():
    per = BoxMesh([6, 5, 4], [0, 0, 0], [1, 1, 1], [True, True, True], "natural")
    assert per.n_interior == 120 and per.n_faces == 360 and per.n_cells == 120
    wall = BoxMesh([6, 5, 4], [0, 0, 0], [1, 1, 1], [False, False, False], "brick")
    assert wall.n_faces == 7 * 5 * 4 + 6 * 6 * 4 + 6 * 5 * 5
    assert wall.n_cells - wall.n_interior == 2 * (5 * 4 + 6 * 4 + 6 * 5)
    # orderings are permutations of the same cell set
    for order in ("natural", "brick", "rcm", "column"):
        m = BoxMesh([6, 5, 4], [0, 0, 0], [1, 1, 1], [True, True, True], order)
        c = np.sort(np.round(m.arrays.cell_centroid[:120] * 1000).view(np.float64).reshape(-1, 3), axis=0)
        np.testing.assert_allclose(c, np.sort(np.round(per.arrays.cell_centroid * 1000), axis=0))


def test_tet_mesh_counts_and_jitter():
    m = TetBoxMesh(5, 0.2, seed=7)
    assert m.n_interior == 6 * 125
    assert m.n_cells - m.n_interior == 6 * 25 * 2  # two triangles per boundary square
    assert np.all(m.arrays.cell_volume > 0)
    m2 = TetBoxMesh(5, 0.2, seed=7)
    np.testing.assert_array_equal(m.arrays.cell_centroid, m2.arrays.cell_centroid)
    with pytest.raises(RuntimeError):
        TetBoxMesh(3, 0.6)


def test_column_order_walks_8x8_columns_layer_by_layer():
    """Column order: 8 x 8 (x, y) columns in x-major order, each walked in z one layer after the other, so every 64
    consecutive cells of a full column are one z-layer and consecutive layers of a column share 64 faces."""
    m = BoxMesh([16, 8, 5], [0, 0, 0], [16, 8, 5], [True, True, True], "column")
    x = m.arrays.cell_centroid[: m.n_interior]
    ijk = np.floor(x).astype(int)
    assert sorted(map(tuple, ijk)) == sorted((i, j, k) for i in range(16) for j in range(8) for k in range(5))
    for c in range(m.n_interior):
        col, rest = divmod(c, 64 * 5)
        layer, pos = divmod(rest, 64)
        assert tuple(ijk[c]) == (8 * col + pos % 8, pos // 8, layer), c
    with pytest.raises(RuntimeError, match="column"):
        BoxMesh([8, 8], [0, 0], [1, 1], [True, True], "column")
