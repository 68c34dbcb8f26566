"""Generated finite-volume meshes (DMPlex-style FV geometry) exposed as numpy views.

``BoxMesh`` mirrors ablate::domain::BoxMesh + GhostBoundaryCells (src/domain/boxMesh.cpp:17-43,
src/domain/modifiers/ghostBoundaryCells.cpp:4-8); ``TetBoxMesh`` is the jittered tetrahedral unit cube;
``GmshMesh`` reads a Gmsh file as ablate::domain::MeshFile + GhostBoundaryCells do (src/domain/meshFile.cpp).
Arrays are views of memory owned by the native mesh object and live as long as the Python object.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class MeshArrays:
    dim: int
    n_cells: int
    n_interior: int
    n_faces: int
    face_cells: np.ndarray  # [F, 2] int32
    face_normal: np.ndarray  # [F, 3]
    face_centroid: np.ndarray  # [F, 3]
    face_shift: np.ndarray | None  # [F, 3]
    cell_centroid: np.ndarray  # [N, 3]
    cell_volume: np.ndarray  # [N]
    face_bc: np.ndarray | None  # [F] int32


def _view(ptr, shape, dtype, owner):
    """numpy view of native memory; the view's base keeps `owner` (the native mesh) alive, so arrays taken from a
    temporary mesh object stay valid."""
    if not ptr:
        return None
    count = int(np.prod(shape))
    ctype = ctypes.c_double if dtype == np.float64 else ctypes.c_int32

    class _Buf(ctype * count):  # a subclass can carry the owner reference
        pass

    buf = _Buf.from_address(ctypes.cast(ptr, ctypes.c_void_p).value)
    buf._owner = owner
    return np.frombuffer(buf, dtype=dtype).reshape(shape)


class _NativeMesh:
    def __init__(self, handle: ctypes.c_void_p):
        self._handle = handle
        desc = _lib.MeshDesc()
        _lib.check(_lib.lib().abl_fv_mesh_describe(handle, ctypes.byref(desc)), None, "abl_fv_mesh_describe")
        N, F = desc.n_cells, desc.n_faces
        self.arrays = MeshArrays(
            dim=desc.dim,
            n_cells=N,
            n_interior=desc.n_interior,
            n_faces=F,
            face_cells=_view(desc.face_cells, (F, 2), np.int32, self),
            face_normal=_view(desc.face_normal, (F, 3), np.float64, self),
            face_centroid=_view(desc.face_centroid, (F, 3), np.float64, self),
            face_shift=_view(desc.face_shift, (F, 3), np.float64, self),
            cell_centroid=_view(desc.cell_centroid, (N, 3), np.float64, self),
            cell_volume=_view(desc.cell_volume, (N,), np.float64, self),
            face_bc=_view(desc.face_bc, (F,), np.int32, self),
        )
        self.desc = desc

    def __del__(self):
        if getattr(self, "_handle", None):
            try:
                _lib.lib().abl_fv_mesh_free(self._handle)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass
            self._handle = None

    # convenience accessors
    @property
    def dim(self):
        return self.arrays.dim

    @property
    def n_cells(self):
        return self.arrays.n_cells

    @property
    def n_interior(self):
        return self.arrays.n_interior

    @property
    def n_faces(self):
        return self.arrays.n_faces

    def source_cells(self) -> np.ndarray:
        """The producer's index of every interior cell, in this mesh's cell order (abl_fv_mesh_source_cells): the
        DMPlex point number, the Gmsh element ordinal, the lexicographic box index (x fastest) or the tet generation
        index.  The same physical cell has the same index under every cell ordering."""
        out = np.empty(self.n_interior, dtype=np.int64)
        _lib.check(_lib.lib().abl_fv_mesh_source_cells(self._handle, out.ctypes.data), None, "abl_fv_mesh_source_cells")
        return out


class ArrayMesh:
    """A mesh given by the caller's own FV geometry arrays (abl_fv_mesh_desc: e.g. the PetscFVFaceGeom /
    PetscFVCellGeom arrays of DMPlexComputeGeometryFVM and the face supports of an ABLATE domain).  The arrays are
    held for the mesh's lifetime; ``face_bc`` may be None when there are no boundary faces."""

    def __init__(self, arrays: MeshArrays, source=None):
        a = arrays
        i32 = lambda x: None if x is None else np.ascontiguousarray(x, dtype=np.int32)
        f64 = lambda x: None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        self.arrays = MeshArrays(dim=int(a.dim), n_cells=int(a.n_cells), n_interior=int(a.n_interior),
                                 n_faces=int(a.n_faces), face_cells=i32(a.face_cells), face_normal=f64(a.face_normal),
                                 face_centroid=f64(a.face_centroid), face_shift=f64(a.face_shift),
                                 cell_centroid=f64(a.cell_centroid), cell_volume=f64(a.cell_volume),
                                 face_bc=i32(a.face_bc))
        m = self.arrays
        ptr = lambda x, t: None if x is None else x.ctypes.data_as(ctypes.POINTER(t))
        self.desc = _lib.MeshDesc(dim=m.dim, n_cells=m.n_cells, n_interior=m.n_interior, n_faces=m.n_faces,
                                  face_cells=ptr(m.face_cells, ctypes.c_int32),
                                  face_normal=ptr(m.face_normal, ctypes.c_double),
                                  face_centroid=ptr(m.face_centroid, ctypes.c_double),
                                  face_shift=ptr(m.face_shift, ctypes.c_double),
                                  cell_centroid=ptr(m.cell_centroid, ctypes.c_double),
                                  cell_volume=ptr(m.cell_volume, ctypes.c_double),
                                  face_bc=ptr(m.face_bc, ctypes.c_int32))
        self._source = np.arange(m.n_interior, dtype=np.int64) if source is None else np.asarray(source, np.int64)

    dim = property(lambda self: self.arrays.dim)
    n_cells = property(lambda self: self.arrays.n_cells)
    n_interior = property(lambda self: self.arrays.n_interior)
    n_faces = property(lambda self: self.arrays.n_faces)

    def source_cells(self) -> np.ndarray:
        return self._source


class BoxMesh(_NativeMesh):
    """Uniform quad/hex box. ``periodic`` per direction; non-periodic sides get ghost cells whose boundary faces
    carry boundary-condition slots 2*d (lower side) and 2*d+1 (upper side)."""

    def __init__(self, faces, lower, upper, periodic=None, ordering: str = "brick"):
        dim = len(faces)
        periodic = periodic if periodic is not None else [True] * dim
        cells = (ctypes.c_int32 * 3)(*[int(x) for x in faces], *([1] * (3 - dim)))
        lo = (ctypes.c_double * 3)(*[float(x) for x in lower], *([0.0] * (3 - dim)))
        hi = (ctypes.c_double * 3)(*[float(x) for x in upper], *([1.0] * (3 - dim)))
        per = (ctypes.c_int32 * 3)(*[int(bool(x)) for x in periodic], *([0] * (3 - dim)))
        h = ctypes.c_void_p()
        _lib.check(
            _lib.lib().abl_fv_mesh_box(dim, cells, lo, hi, per, _lib.ORDER[ordering], ctypes.byref(h)),
            None,
            "abl_fv_mesh_box",
        )
        self.faces, self.lower, self.upper, self.periodic = list(faces), list(lower), list(upper), list(periodic)
        super().__init__(h)


class TetBoxMesh(_NativeMesh):
    """Unit cube of n^3 cubes x 6 Kuhn tetrahedra, interior nodes jittered by up to ``jitter`` * h."""

    def __init__(self, n: int, jitter: float = 0.2, seed: int = 20261202, ordering: str = "rcm"):
        h = ctypes.c_void_p()
        _lib.check(
            _lib.lib().abl_fv_mesh_tet_box(int(n), float(jitter), int(seed), _lib.ORDER[ordering], ctypes.byref(h)),
            None,
            "abl_fv_mesh_tet_box",
        )
        self.n, self.jitter, self.seed = n, jitter, seed
        super().__init__(h)


def gmsh_physical_names(path) -> dict:
    """The $PhysicalNames section of a Gmsh file: name -> (dimension, tag)."""
    names, lines = {}, open(path, encoding="utf-8", errors="replace").read().splitlines()
    if "$PhysicalNames" not in lines:
        return names
    i = lines.index("$PhysicalNames")
    for line in lines[i + 2 : i + 2 + int(lines[i + 1])]:
        dim, tag, name = line.split(maxsplit=2)
        names[name.strip().strip('"')] = (int(dim), int(tag))
    return names


class GmshMesh(_NativeMesh):
    """Gmsh MSH 4.1 ASCII mesh with one ghost cell per boundary face; boundary faces carry their physical tag as
    boundary-condition slot (the label value of a boundary condition's labelIds).  ``labels`` maps the physical
    names of the boundary groups to those slots."""

    def __init__(self, path, ordering: str = "natural"):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().abl_fv_mesh_gmsh(str(path).encode(), _lib.ORDER[ordering], ctypes.byref(h)), None,
                   "abl_fv_mesh_gmsh")
        self.path = str(path)
        super().__init__(h)
        bdim = self.arrays.dim - 1
        self.labels = {n: t for n, (d, t) in gmsh_physical_names(path).items() if d == bdim}



class DMPlexMesh(_NativeMesh):
    """Mesh from a DMPlex topology (abl_fv_mesh_dmplex): ``cones`` the cone of every point (a list of point lists, or
    ``(cone_offsets, cone_points)``), ``vertex_coords`` [nv][dim] the depth-0 points' coordinates in point order,
    ``face_sets`` {value: face points} of the "Face Sets" label."""

    def __init__(self, dim: int, cones, vertex_coords, face_sets=None, ordering: str = "natural"):
        if isinstance(cones, tuple):
            off, pts = (np.ascontiguousarray(x, dtype=np.int64) for x in cones)
        else:
            off = np.concatenate([[0], np.cumsum([len(c) for c in cones])]).astype(np.int64)
            pts = np.array([q for c in cones for q in c], dtype=np.int64)
        n = off.size - 1
        X = np.ascontiguousarray(vertex_coords, dtype=np.float64)
        fs = None
        if face_sets:
            fs = np.full(n, -1, dtype=np.int32)
            for v, idx in face_sets.items():
                fs[np.asarray(idx, dtype=np.int64)] = int(v)
        h = ctypes.c_void_p()
        _lib.check(
            _lib.lib().abl_fv_mesh_dmplex(int(dim), n, off.ctypes.data, pts.ctypes.data, X.ctypes.data,
                                          None if fs is None else fs.ctypes.data, _lib.ORDER[ordering],
                                          ctypes.byref(h)),
            None,
            "abl_fv_mesh_dmplex",
        )
        super().__init__(h)
        self.labels = {int(v): int(v) for v in (face_sets or {})}


class Hdf5Mesh(DMPlexMesh):
    """The mesh of an ABLATE Hdf5Serializer output file (src/io/hdf5Serializer.cpp:23-160, PETSc's
    DMPlexView_HDF5): /topology/order (global point number of each stored point), /topology/cones (cone sizes),
    /topology/cells (the cones, concatenated, in global point numbers), /geometry/vertices (rows in vertex point
    order) and /labels/<label>/<value>/indices.  ``fields(name)`` returns a saved field of the last time step as
    [n_comp, n_interior] in this mesh's cell order."""

    def __init__(self, path, ordering: str = "natural", label: str = "Face Sets"):
        from .hdf5 import H5File

        h = H5File(path)
        order = h.read("/topology/order").ravel().astype(np.int64)
        csize = h.read("/topology/cones").ravel().astype(np.int64)
        flat = h.read("/topology/cells").ravel().astype(np.int64)
        n = order.size
        if np.any(np.sort(order) != np.arange(n)):
            raise ValueError(f"{path}: /topology/order is not a permutation of the points")
        # re-index the stored cones by global point number
        start = np.concatenate([[0], np.cumsum(csize)])
        size_g = np.empty(n, dtype=np.int64)
        size_g[order] = csize
        off = np.concatenate([[0], np.cumsum(size_g)]).astype(np.int64)
        # element i of stored cone p lands at off[order[p]] + i
        dst = np.repeat(off[order] - start[:-1], csize) + np.arange(flat.size, dtype=np.int64)
        pts = np.empty(off[-1], dtype=np.int64)
        pts[dst] = flat
        X = h.read("/geometry/vertices")
        dim = X.shape[1]
        face_sets = {}
        if h.has(f"/labels/{label}"):
            for v in h.keys(f"/labels/{label}"):
                face_sets[int(v)] = h.read(f"/labels/{label}/{v}/indices").ravel().astype(np.int64)
        self.h5, self.path = h, str(path)
        super().__init__(dim, (off, pts), X, face_sets, ordering)
        self._cstart = int(self.source_cells().min())

    def fields(self, name: str = "solution", group: str = "fields") -> np.ndarray:
        """Field `name` of the last saved time step as [n_comp, n_interior] in this mesh's cell order (the file
        stores it per cell in point order from the first cell point on)."""
        data = self.h5.read(f"/{group}/{name}")
        last = data[-1] if data.ndim > 1 else data
        last = last.reshape(last.shape[0], -1)
        rows = self.source_cells() - self._cstart
        return np.ascontiguousarray(last[rows].T)
