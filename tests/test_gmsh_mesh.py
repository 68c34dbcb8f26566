"""Gmsh ingestion (abl_fv_mesh_gmsh) on the reference's own mesh files (tests/golden/*.msh, copied data) and on
small 3-D files written here, checked against what the geometry fixes: total volume, first moments, closure of
every cell (sum of outward area normals = 0), boundary normals summing to the area vectors of the domain's sides,
and ghost cells mirroring their interior cell."""
from pathlib import Path

import numpy as np
import pytest

from ablate_amd.mesh import GmshMesh

GOLDEN = Path(__file__).resolve().parent / "golden"
PIPE = GOLDEN / "pipeFlow.msh"  # tests/integrationTests/inputs/compressibleFlow/gmshPipeFlow/pipeFlow.msh
MIXED = GOLDEN / "mixedCells2D.msh"  # tests/integrationTests/inputs/compressibleFlow/mixedCellTypeTest/


def closure(a):
    S = np.zeros((a.n_cells, 3))
    np.add.at(S, a.face_cells[:, 0], a.face_normal)
    np.add.at(S, a.face_cells[:, 1], -a.face_normal)
    return np.abs(S[: a.n_interior]).max()


def label_normal_sum(a, tag):
    return a.face_normal[a.face_bc == tag].sum(axis=0)


@pytest.mark.parametrize("ordering", ["natural", "rcm", "morton"])
def test_reference_pipe_mesh(ordering):
    """0.5 x 0.1 channel: inlet at x = 0, outlet at x = 0.5, walls at y = 0 and 0.1."""
    m = GmshMesh(PIPE, ordering)
    a = m.arrays
    NI = a.n_interior
    assert a.dim == 2 and m.labels == {"inlet": 5, "wall": 6, "outlet": 7}
    assert a.n_cells - NI == int((a.face_bc >= 0).sum())
    assert a.cell_volume[:NI].sum() == pytest.approx(0.05, rel=1e-13)
    moment = (a.cell_volume[:NI, None] * a.cell_centroid[:NI]).sum(axis=0)
    assert moment[:2] == pytest.approx([0.25 * 0.05, 0.05 * 0.05], rel=1e-12)
    assert closure(a) < 1e-17
    assert label_normal_sum(a, 5) == pytest.approx([-0.1, 0, 0], abs=1e-15)
    assert label_normal_sum(a, 7) == pytest.approx([0.1, 0, 0], abs=1e-15)
    assert label_normal_sum(a, 6) == pytest.approx([0, 0, 0], abs=1e-15)
    # every boundary face points out of the domain and its ghost mirrors the interior cell
    for f in np.nonzero(a.face_bc >= 0)[0][:20]:
        L, G = a.face_cells[f]
        assert G >= NI and L < NI
        n = a.face_normal[f]
        assert np.dot(n, a.face_centroid[f] - a.cell_centroid[L]) > 0
        mid = 0.5 * (a.cell_centroid[L] + a.cell_centroid[G])
        assert np.dot(n, mid - a.face_centroid[f]) == pytest.approx(0, abs=1e-15)
        assert a.cell_volume[G] == a.cell_volume[L]


def test_reference_mixed_cell_mesh():
    """Triangles and quadrilaterals on 0.03 x 0.01."""
    m = GmshMesh(MIXED)
    a = m.arrays
    NI = a.n_interior
    assert m.labels == {"inlet": 11, "outlet": 12, "wall": 13}
    assert a.cell_volume[:NI].sum() == pytest.approx(3e-4, rel=1e-13)
    moment = (a.cell_volume[:NI, None] * a.cell_centroid[:NI]).sum(axis=0)
    assert moment[:2] == pytest.approx([0.015 * 3e-4, 0.005 * 3e-4], rel=1e-12)
    assert closure(a) < 1e-18
    assert label_normal_sum(a, 11) == pytest.approx([-0.01, 0, 0], abs=1e-16)
    assert label_normal_sum(a, 12) == pytest.approx([0.01, 0, 0], abs=1e-16)
    faces_per_cell = np.bincount(a.face_cells[:, 0], minlength=a.n_cells) + np.bincount(a.face_cells[:, 1],
                                                                                          minlength=a.n_cells)
    assert set(np.unique(faces_per_cell[:NI])) == {3, 4}  # both element kinds present


def write_msh(path, nodes, cells, cell_type, bfaces, bdim, tag=1):
    """A minimal MSH 4.1 ASCII file: one entity holding the cells, one boundary entity (physical `tag`) holding the
    boundary elements (lines, triangles or quadrangles by node count)."""
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", "1", f'{bdim} {tag} "walls"',
             "$EndPhysicalNames", "$Entities"]
    counts = [0, 0, 0, 0]
    counts[bdim] += 1
    counts[bdim + 1] += 1
    lines.append(" ".join(map(str, counts)))
    for d in (bdim, bdim + 1):  # entity tags: boundary bdim + 1, cells bdim + 2
        phys = f"1 {tag}" if d == bdim else "0"
        lines.append(f"{d + 1} 0 0 0 1 1 1 {phys} 0")
    lines += ["$EndEntities", "$Nodes", f"1 {len(nodes)} 1 {len(nodes)}", f"{bdim + 1} {bdim + 2} 0 {len(nodes)}"]
    lines += [str(i + 1) for i in range(len(nodes))]
    lines += [" ".join(f"{x:.17g}" for x in p) for p in nodes]
    groups = {}
    for fc in bfaces:
        groups.setdefault({2: 1, 3: 2, 4: 3}[len(fc)], []).append(fc)
    n_el = len(cells) + len(bfaces)
    lines += ["$EndNodes", "$Elements", f"{len(groups) + 1} {n_el} 1 {n_el}"]
    k = 1
    for etype, fcs in groups.items():
        lines.append(f"{bdim} {bdim + 1} {etype} {len(fcs)}")
        for fc in fcs:
            lines.append(" ".join(map(str, [k] + [v + 1 for v in fc])))
            k += 1
    lines.append(f"{bdim + 1} {bdim + 2} {cell_type} {len(cells)}")
    for c in cells:
        lines.append(" ".join(map(str, [k] + [v + 1 for v in c])))
        k += 1
    lines.append("$EndElements")
    Path(path).write_text("\n".join(lines) + "\n")


def _hex_grid(n, skew=0.0):
    """n^3 hexahedra on the unit cube (Gmsh node order), x-nodes sheared by `skew` * z (planar faces kept)."""
    idx = lambda i, j, k: i + (n + 1) * (j + (n + 1) * k)
    nodes = [(i / n + skew * (k / n), j / n, k / n) for k in range(n + 1) for j in range(n + 1) for i in range(n + 1)]
    cells, quads = [], []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                cells.append([idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k),
                              idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1), idx(i, j + 1, k + 1)])
    for a in range(n):
        for b in range(n):
            quads += [[idx(0, a, b), idx(0, a + 1, b), idx(0, a + 1, b + 1), idx(0, a, b + 1)],
                      [idx(n, a, b), idx(n, a + 1, b), idx(n, a + 1, b + 1), idx(n, a, b + 1)],
                      [idx(a, 0, b), idx(a + 1, 0, b), idx(a + 1, 0, b + 1), idx(a, 0, b + 1)],
                      [idx(a, n, b), idx(a + 1, n, b), idx(a + 1, n, b + 1), idx(a, n, b + 1)],
                      [idx(a, b, 0), idx(a + 1, b, 0), idx(a + 1, b + 1, 0), idx(a, b + 1, 0)],
                      [idx(a, b, n), idx(a + 1, b, n), idx(a + 1, b + 1, n), idx(a, b + 1, n)]]
    return nodes, cells, quads


def test_hexahedra_and_prisms(tmp_path):
    """A sheared 3^3 hex block (volume 1, centroid shifted by skew / 2 in x), then the same block with each hex cut
    into two prisms."""
    n, skew = 3, 0.3
    nodes, hexes, quads = _hex_grid(n, skew)
    p = tmp_path / "hex.msh"
    write_msh(p, nodes, hexes, 5, quads, 2)
    a = GmshMesh(p).arrays
    NI = a.n_interior
    assert NI == 27 and a.n_cells - NI == 54
    assert a.cell_volume[:NI].sum() == pytest.approx(1.0, rel=1e-13)
    assert a.cell_volume[:NI] == pytest.approx(np.full(NI, 1 / 27), rel=1e-12)
    moment = (a.cell_volume[:NI, None] * a.cell_centroid[:NI]).sum(axis=0)
    assert moment == pytest.approx([0.5 + skew / 2, 0.5, 0.5], rel=1e-12)
    assert closure(a) < 1e-15
    assert label_normal_sum(a, 1) == pytest.approx([0, 0, 0], abs=1e-14)
    # prisms: split every hex along its (0, 2, 6, 4) diagonal plane
    prisms = []
    for h in hexes:
        prisms += [[h[0], h[1], h[2], h[4], h[5], h[6]], [h[0], h[2], h[3], h[4], h[6], h[7]]]
    # the z = 0 and z = 1 quads split along the same diagonal; the side quads stay
    bnd = []
    for q in quads:
        zs = {round(nodes[v][2], 12) for v in q}
        bnd += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]] if len(zs) == 1 else [q]
    p2 = tmp_path / "prism.msh"
    write_msh(p2, nodes, prisms, 6, bnd, 2)
    b = GmshMesh(p2).arrays
    assert b.n_interior == 54
    assert b.cell_volume[: b.n_interior].sum() == pytest.approx(1.0, rel=1e-13)
    assert closure(b) < 1e-15


def test_tetrahedra_and_pyramids(tmp_path):
    """The unit cube as six pyramids around its centre (apex at the centre, one per face), and as twelve tets
    (each pyramid split in two): volumes 1/6 and 1/12, centroids a quarter of the way from base to apex."""
    c = (0.5, 0.5, 0.5)
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    nodes = corners + [c]
    faces = [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [2, 3, 7, 6], [0, 4, 7, 3], [1, 2, 6, 5]]
    pyramids = [f + [8] for f in faces]
    p = tmp_path / "pyr.msh"
    write_msh(p, nodes, pyramids, 7, faces, 2)
    a = GmshMesh(p).arrays
    assert a.n_interior == 6
    assert a.cell_volume[:6] == pytest.approx(np.full(6, 1 / 6), rel=1e-13)
    assert closure(a) < 1e-15
    for k in range(6):
        # centroid of a pyramid: 1/4 of the way from the base centroid to the apex
        d = a.cell_centroid[k] - np.array(c)
        assert np.linalg.norm(d) == pytest.approx(0.375, rel=1e-12)
    tets, tris = [], []
    for f in faces:
        tets += [[f[0], f[1], f[2], 8], [f[0], f[2], f[3], 8]]
        tris += [[f[0], f[1], f[2]], [f[0], f[2], f[3]]]
    p2 = tmp_path / "tet.msh"
    write_msh(p2, nodes, tets, 4, tris, 2)
    b = GmshMesh(p2).arrays
    assert b.n_interior == 12 and b.n_cells == 24
    assert b.cell_volume[:12] == pytest.approx(np.full(12, 1 / 12), rel=1e-13)
    assert closure(b) < 1e-15


def test_rejects_unsupported_files(tmp_path):
    p = tmp_path / "old.msh"
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
    with pytest.raises(RuntimeError, match="4.1"):
        GmshMesh(p)
    with pytest.raises(RuntimeError, match="cannot open"):
        GmshMesh(tmp_path / "missing.msh")


def test_morton_order_is_a_local_permutation():
    """Morton order of a tet mesh: the same cells (volumes and centroids as multisets), faces numbered by their left
    cell, and neighbours close in index (the point of the order: median |R - L| far below the RCM order's)."""
    from ablate_amd.mesh import TetBoxMesh

    a, b = TetBoxMesh(12, 0.2, ordering="rcm").arrays, TetBoxMesh(12, 0.2, ordering="morton").arrays
    NI = a.n_interior
    assert np.allclose(np.sort(a.cell_volume[:NI]), np.sort(b.cell_volume[:NI]), rtol=0, atol=0)
    key = lambda x: np.lexsort(x[:NI].T)
    assert np.array_equal(a.cell_centroid[:NI][key(a.cell_centroid)], b.cell_centroid[:NI][key(b.cell_centroid)])
    assert np.all(np.diff(b.face_cells[:, 0]) >= 0)
    gap = lambda m: np.median(np.abs(np.diff(m.face_cells[m.face_cells[:, 1] < NI], axis=1)))
    assert gap(b) * 20 < gap(a)
