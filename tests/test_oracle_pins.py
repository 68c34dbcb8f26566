"""Pins of the oracle's build-defined and third-party parts by what the mathematics fixes, each with a planted
mistake (an oracle variant with one formula wrong) that the pin rejects:

  - Rusanov (build-defined; Rusanov 1961, Toro 2009 sec. 10.5): consistency with the physical flux, antisymmetry,
    and the closed form of a density jump at rest.
  - Barth-Jespersen (build-defined; Barth & Jespersen, AIAA-89-0366): the limited gradient is the largest multiple
    phi in [0, 1] of the least-squares gradient whose face projections stay within the neighbour min / max
    (bounded, and tight when phi < 1); linear data is left unlimited on a uniform mesh.
  - PETSc's symmetric-form limiters (PetscLimiterLimit_*): phi(f) = 2 psi(r) / (1 + r) with r = f / (1 - f) for the
    textbook TVD limiters psi(r) (minmod; van Leer 1974; van Albada 1982; Roe's superbee; van Leer's MC, 1977).
  - Second-order assembly (cellInterpolant.cpp:330-639): for a state linear in x the least-squares gradients are
    exact, both face states equal the field at the face centroid, and the residual of every cell without a ghost
    face is -(1/V) sum_f +/- F(U(x_f)).n_f of the physical Euler flux.
  - FaceInterpolant (faceInterpolant.cpp:109-249): face values and gradients of linear fields exact at every face
    with two supports (faceInterpolantTests.cpp:224-265: 5^d boxes, 1e-8).
  - Periodic wrap faces: translation invariance; a periodic box's second-order residual equals that of the same
    cells inside a box three times larger, where no stencil reaches a wrap face.
"""
import numpy as np
import pytest

from ablate_amd.mesh import BoxMesh, MeshArrays, TetBoxMesh
from oracle import binding as orc

G, RG = 1.4, 287.0


def physical_flux(U, n):
    """Euler flux F(U).n of conserved states U [nc, m] (rho, rhoE, rhoU..) through normals n [m, dim]."""
    dim = n.shape[1]
    rho = U[0]
    vel = U[2 : 2 + dim] / rho
    p = (G - 1.0) * (U[1] - 0.5 * rho * np.sum(vel * vel, axis=0))
    un = np.sum(vel * n.T, axis=0)
    F = np.empty_like(U[: dim + 2])
    F[0] = rho * un
    F[1] = (U[1] + p) * un
    F[2:] = U[2 : 2 + dim] * un + p * n.T
    return F


def _state(rho, vel, p):
    vel = np.asarray(vel, dtype=np.float64)
    return np.array([rho, p / (G - 1.0) + 0.5 * rho * vel @ vel, *(rho * vel)])


# ------------------------------------------------------------------------------------------- Rusanov
def _rusanov_checks():
    rng = np.random.default_rng(4)
    bad = 0
    for _ in range(40):
        n = rng.normal(size=3) * rng.uniform(0.2, 2.0)
        A = np.linalg.norm(n)
        uL = _state(rng.uniform(0.5, 2), rng.uniform(-200, 200, 3), rng.uniform(5e4, 2e5))
        uR = _state(rng.uniform(0.5, 2), rng.uniform(-200, 200, 3), rng.uniform(5e4, 2e5))
        # consistency: equal states give the physical flux through the face
        f = orc.rusanov_flux(n, uL, uL)
        bad += not np.allclose(f, physical_flux(uL[:, None], n[None, :] / A)[:, 0] * A, rtol=1e-12, atol=1e-9)
        # antisymmetry under side swap with the normal reversed
        bad += not np.allclose(orc.rusanov_flux(n, uL, uR), -orc.rusanov_flux(-n, uR, uL), rtol=1e-12, atol=1e-9)
        # density jump at rest, equal pressure: F_rho = -1/2 lambda (rho_R - rho_L) A with lambda = max(a_L, a_R),
        # F_rhoE = 0, F_rhoU = p n
        p = rng.uniform(5e4, 2e5)
        rl, rr = rng.uniform(0.3, 3, 2)
        f = orc.rusanov_flux(n, _state(rl, [0, 0, 0], p), _state(rr, [0, 0, 0], p))
        lam = max(np.sqrt(G * p / rl), np.sqrt(G * p / rr))
        want = np.concatenate([[-0.5 * lam * (rr - rl) * A, 0.0], p * n])
        bad += not np.allclose(f, want, rtol=1e-12, atol=1e-9 * p * A)
    return bad


def test_rusanov_closed_forms():
    assert _rusanov_checks() == 0


def test_rusanov_species_closed_form():
    """Species ride the same flux: F_sp = A [1/2 (rhoY_L un_L + rhoY_R un_R) - 1/2 lambda (rhoY_R - rhoY_L)]; with
    equal velocities and sound speeds the rhoY flux is Y times the density flux for a common Y."""
    n = np.array([0.3, -0.2, 0.9])
    uL = np.concatenate([_state(1.1, [30, 5, -3], 1.0e5), [0.33, 0.77]])
    uR = np.concatenate([_state(1.1, [30, 5, -3], 1.2e5), [0.33, 0.77]])
    f = orc.rusanov_flux(n, uL, uR, n_species=2)
    np.testing.assert_allclose(f[5:], f[0] * np.array([0.3, 0.7]), rtol=1e-12)


def test_rusanov_planted_mistake_fails():
    with orc.planted("rusanov_min_speed"):
        assert _rusanov_checks() > 10


# ------------------------------------------------------------------------------------------- Barth-Jespersen
def _faces_of(a):
    cells = [[] for _ in range(a.n_interior)]
    for f, (l, r) in enumerate(a.face_cells):
        if l < a.n_interior:
            cells[l].append((f, r))
        if r < a.n_interior:
            cells[r].append((f, l))
    return cells


def _bj_violations(mesh, U):
    """Counts cells / components that break the Barth-Jespersen characterisation."""
    a = mesh.arrays
    NI = a.n_interior
    ns = U.shape[0] - a.dim - 2
    ph_bj = orc.physics(n_species=ns, recon="leastsquares", limiter="barthjespersen")
    ph_no = orc.physics(n_species=ns, recon="leastsquares", limiter="none")
    bcs = {s: {"kind": orc.BC_OUTFLOW} for s in range(6)}
    _, Uf = orc.rhs(a, ph_no, U.copy(), bcs)  # ghost rows filled (outflow copies)
    Gl = orc.cell_gradients(a, ph_bj, Uf)
    G0 = orc.cell_gradients(a, ph_no, Uf)
    faces = _faces_of(a)
    bad = 0
    for c in range(NI):
        xc = a.cell_centroid[c]
        for p in range(U.shape[0]):
            g0, g = G0[c, p], Gl[c, p]
            n0 = g0 @ g0
            phi = (g @ g0) / n0 if n0 > 0 else 1.0
            if not (-1e-15 <= phi <= 1 + 1e-15) or np.abs(g - phi * g0).max() > 1e-12 * np.abs(g0).max():
                bad += 1
                continue
            u = Uf[p, c]
            nb = np.array([Uf[p, n] for _, n in faces[c]] + [u])
            lo, hi = nb.min(), nb.max()
            proj = np.array([u + g @ (a.face_centroid[f] - xc) for f, _ in faces[c]])
            tol = 1e-12 * max(abs(u), 1e-300)
            if proj.min() < lo - tol or proj.max() > hi + tol:
                bad += 1  # a limited face value outside the neighbour range
            elif phi < 1 - 1e-12 and not (abs(proj.max() - hi) <= tol or abs(proj.min() - lo) <= tol):
                bad += 1  # not the largest admissible phi
    return bad


def _bj_state(mesh):
    from tests.fvcases import Case

    return Case("bj", mesh, n_species=2, mach=0.6).initial_state()


@pytest.mark.parametrize("make", [lambda: TetBoxMesh(3, 0.2, ordering="natural"),
                                  lambda: BoxMesh([5, 4, 3], [0, 0, 0], [1, 1, 1], [False] * 3, ordering="natural")],
                         ids=["tets", "hex"])
def test_barth_jespersen_is_the_largest_bounded_multiple(make):
    mesh = make()
    assert _bj_violations(mesh, _bj_state(mesh)) == 0


def test_barth_jespersen_planted_mistake_fails():
    mesh = TetBoxMesh(3, 0.2, ordering="natural")
    U = _bj_state(mesh)
    with orc.planted("bj_neighbor"):
        assert _bj_violations(mesh, U) > 10


def test_barth_jespersen_leaves_linear_data_unlimited_on_a_uniform_mesh():
    mesh = BoxMesh([6, 5, 4], [0, 0, 0], [1, 1, 1], [True] * 3, ordering="natural")
    a = mesh.arrays
    NI = mesh.n_interior
    x = a.cell_centroid[:NI]
    # periodic linear data needs a periodic-compatible field: use a field linear inside the box interior cells and
    # check only cells whose neighbours do not wrap
    U = np.zeros((5, mesh.n_cells))
    U[0, :NI] = 1.0 + 0.1 * x[:, 0] - 0.05 * x[:, 1] + 0.02 * x[:, 2]
    U[1, :NI] = 2.5e5 + 1e3 * x[:, 1]
    U[2, :NI] = 3.0 + x[:, 2]
    U[3, :NI] = -2.0 + 0.5 * x[:, 0]
    U[4, :NI] = 1.0 - x[:, 1]
    Gl = orc.cell_gradients(a, orc.physics(recon="leastsquares", limiter="barthjespersen"), U)
    G0 = orc.cell_gradients(a, orc.physics(recon="leastsquares", limiter="none"), U)
    h = np.array([1 / 6, 1 / 5, 1 / 4])
    inner = np.all((x > 1.5 * h) & (x < 1 - 1.5 * h), axis=1)
    np.testing.assert_array_equal(Gl[:NI][inner], G0[:NI][inner])
    np.testing.assert_allclose(G0[:NI][inner][:, 0], np.tile([0.1, -0.05, 0.02], (inner.sum(), 1)), rtol=1e-12)


# ------------------------------------------------------------------------------------------- PETSc limiters
def _sweby(name, r):
    if name == "minmod":
        return max(0.0, min(1.0, r))
    if name == "vanleer":
        return (r + abs(r)) / (1.0 + abs(r))
    if name == "vanalbada":
        return max(0.0, (r * r + r) / (r * r + 1.0))
    if name == "superbee":
        return max(0.0, min(2.0 * r, 1.0), min(r, 2.0))
    if name == "mc":
        return max(0.0, min(2.0 * r, 0.5 * (1.0 + r), 2.0))
    raise KeyError(name)


def _limiter_mismatches(name):
    bad = 0
    for f in np.linspace(0.01, 0.99, 99):
        r = f / (1.0 - f)
        want = 2.0 * _sweby(name, r) / (1.0 + r)
        bad += abs(orc.limiter(name, f) - want) > 1e-13
    return bad


@pytest.mark.parametrize("name", ["minmod", "vanleer", "vanalbada", "superbee", "mc"])
def test_symmetric_limiters_are_the_textbook_tvd_limiters(name):
    """cellInterpolant.cpp:316-324 evaluates phi(f) with f = 1/2 (u_n - u_c) / (grad u_c . (x_n - x_c)); with
    r = f / (1 - f) the symmetric forms are 2 psi(r) / (1 + r) of the classical limiter functions."""
    assert _limiter_mismatches(name) == 0


def test_limiter_planted_mistake_fails():
    with orc.planted("minmod_as_mc"):
        assert _limiter_mismatches("minmod") > 20


def test_sin_limiter_shape():
    """The sin limiter (PETSc's default, singled out by the reference's MMS table) is the half sine on [0, 1] and
    zero outside: its mean over [0, 1] is 2 / pi and it is not any of the TVD forms above."""
    f = (np.arange(20000) + 0.5) / 20000
    vals = np.array([orc.limiter("sin", x) for x in f])
    assert vals.mean() == pytest.approx(2.0 / np.pi, rel=1e-8)
    assert abs(orc.limiter("sin", 1.3)) < 1e-15 and orc.limiter("sin", -0.1) == 0.0  # sin(pi min(f, 1)), f > 1
    for name in ("minmod", "vanleer", "vanalbada", "superbee", "mc"):
        assert max(abs(orc.limiter(name, x) - orc.limiter("sin", x)) for x in f[::500]) > 0.05


# ------------------------------------------------------------------------------------------- linear exactness
def linear_state(mesh, dim, rng):
    """A conserved state linear in x (physically admissible over the unit box) at interior and ghost centroids."""
    a = mesh.arrays
    U0 = _state(1.2, [40.0, -25.0, 10.0][:dim], 1.1e5)
    Gr = np.zeros((dim + 2, dim))
    Gr[0] = rng.uniform(-0.1, 0.1, dim)
    Gr[1] = rng.uniform(-5e3, 5e3, dim)
    Gr[2:] = rng.uniform(-8.0, 8.0, (dim, dim))
    x = a.cell_centroid[:, :dim]
    return U0[:, None] + Gr @ x.T


def _exact_at(U, a, NI, dim, x):
    """The linear field through the interior cell states evaluated at x (least squares over the interior cells)."""
    X = np.hstack([np.ones((NI, 1)), a.cell_centroid[:NI, :dim]])
    coef, *_ = np.linalg.lstsq(X, U[:, :NI].T, rcond=None)
    return (np.hstack([np.ones((x.shape[0], 1)), x]) @ coef).T


def linear_expected(mesh, U):
    """-(1/V) sum_f +/- F(U(x_f)).n_f of the physical flux at the face centroids, and the cells with no ghost face
    (where the second-order RHS must reproduce it)."""
    a = mesh.arrays
    NI, dim = mesh.n_interior, mesh.dim
    xf = a.face_centroid[:, :dim]
    F = physical_flux(_exact_at(U, a, NI, dim, xf), a.face_normal[:, :dim])
    Rexp = np.zeros_like(U)
    L, Rc = a.face_cells[:, 0], a.face_cells[:, 1]
    np.add.at(Rexp.T, L, -F.T)
    np.add.at(Rexp.T, Rc, F.T)
    Rexp /= a.cell_volume[None, :]
    touched = np.zeros(a.n_cells, bool)
    touched[L[Rc >= NI]] = True
    return Rexp, ~touched[:NI]


def linear_deviation(mesh, U, R):
    Rexp, sel = linear_expected(mesh, U)
    NI = mesh.n_interior
    scale = np.abs(Rexp[:, :NI][:, sel]).max()
    return np.abs(R[:, :NI][:, sel] - Rexp[:, :NI][:, sel]).max() / scale, int(sel.sum())


LINEAR_MESHES = {
    "hex": lambda: BoxMesh([6, 5, 4], [0, 0, 0], [1, 1, 1], [False] * 3, ordering="natural"),
    "tets": lambda: TetBoxMesh(4, 0.2, ordering="natural"),
    "quad": lambda: BoxMesh([7, 6], [0, 0], [1, 1], [False, False], ordering="natural"),
}
LINEAR_CASES = [("hex", "ausm", "none"), ("hex", "ausm", "barthjespersen"), ("hex", "average", "none"),
                ("tets", "rusanov", "none"), ("tets", "ausm", "none"), ("quad", "ausm", "barthjespersen")]


def _linear_deviation(mesh, flux, limiter, seed=3):
    U = linear_state(mesh, mesh.dim, np.random.default_rng(seed))
    bcs = {s: {"kind": orc.BC_GHOST} for s in range(2 * mesh.dim)}
    R, _ = orc.rhs(mesh.arrays, orc.physics(flux=flux, recon="leastsquares", limiter=limiter), U.copy(), bcs)
    return linear_deviation(mesh, U, R)


@pytest.mark.parametrize("mesh_kind,flux,limiter", LINEAR_CASES)
def test_second_order_rhs_is_exact_for_linear_states(mesh_kind, flux, limiter):
    mesh = LINEAR_MESHES[mesh_kind]()
    dev, ncells = _linear_deviation(mesh, flux, limiter)
    assert ncells > 0
    assert dev < 1e-10, dev


def test_linear_exactness_planted_mistake_fails():
    mesh = TetBoxMesh(4, 0.2, ordering="natural")
    with orc.planted("project_sign"):
        dev, _ = _linear_deviation(mesh, "rusanov", "none")
    assert dev > 1e-3


# ------------------------------------------------------------------------------------------- face interpolant
def strip_ghosts(a) -> MeshArrays:
    """The same box without ghost cells and boundary faces (a BoxMesh without the GhostBoundaryCells modifier, as
    faceInterpolantTests.cpp:63-73 builds it)."""
    keep = a.face_cells[:, 1] < a.n_interior
    NI = a.n_interior
    return MeshArrays(dim=a.dim, n_cells=NI, n_interior=NI, n_faces=int(keep.sum()),
                      face_cells=np.ascontiguousarray(a.face_cells[keep]),
                      face_normal=np.ascontiguousarray(a.face_normal[keep]),
                      face_centroid=np.ascontiguousarray(a.face_centroid[keep]), face_shift=None,
                      cell_centroid=np.ascontiguousarray(a.cell_centroid[:NI]),
                      cell_volume=np.ascontiguousarray(a.cell_volume[:NI]), face_bc=None)


def face_test_state(x, dim):
    """rho = 1; velocity and temperature linear; rhoE = rho (cv T + |u|^2 / 2): every aux field is linear, and the
    solution fields rho, rhoU are linear.  Returns U [dim + 2, n] and the exact aux gradients [1 + dim][dim]."""
    cv = RG / (G - 1.0)
    gT = np.array([3.0, -1.0, 2.0])[:dim]
    gU = np.array([[10.0, 3.0, 2.0], [-1.0, -1.0, -1.0], [-2.0, 0.0, 0.0]])[:dim, :dim]
    T = 300.0 + x @ gT
    vel = np.array([20.0, -5.0, 7.0])[:dim, None] + gU @ x.T
    U = np.empty((dim + 2, x.shape[0]))
    U[0] = 1.0
    U[1] = cv * T + 0.5 * np.sum(vel * vel, axis=0)
    U[2:] = vel
    return U, np.vstack([gT[None, :], gU])


@pytest.mark.parametrize("dim", [2, 3])
def test_face_interpolant_exact_for_linear_fields(dim):
    """faceInterpolantTests.cpp:224-265: on a 5^d box without ghost cells every face with two supports gets the exact
    value and gradient of linear fields (1e-8)."""
    box = BoxMesh([5] * dim, [0.0] * dim, [1.0] * dim, [False] * dim, ordering="natural")
    a = strip_ghosts(box.arrays)
    U, gaux = face_test_state(a.cell_centroid[:, :dim], dim)
    fs, fsg, fa, fag = orc.face_interpolate(a, orc.physics(), U)
    xf = a.face_centroid[:, :dim]
    Ue, _ = face_test_state(xf, dim)
    for p in [0] + list(range(2, dim + 2)):  # the linear solution fields: rho, rhoU
        np.testing.assert_allclose(fs[:, p], Ue[p], atol=1e-8)
    np.testing.assert_allclose(fsg[:, 0], 0.0, atol=1e-8)
    np.testing.assert_allclose(fsg[:, 2:], np.broadcast_to(gaux[1:], fsg[:, 2:].shape), atol=1e-8)
    cv = RG / (G - 1.0)
    T = (Ue[1] - 0.5 * np.sum(Ue[2:] ** 2, axis=0)) / cv
    np.testing.assert_allclose(fa[:, 0], T, atol=1e-8)
    np.testing.assert_allclose(fa[:, 1:], Ue[2:].T, atol=1e-8)
    np.testing.assert_allclose(fag, np.broadcast_to(gaux, fag.shape), atol=1e-8)


def test_face_interpolant_planted_mistake_fails():
    box = BoxMesh([5, 5], [0, 0], [1, 1], [False, False], ordering="natural")
    a = strip_ghosts(box.arrays)
    U, _ = face_test_state(a.cell_centroid[:, :2], 2)
    with orc.planted("face_value_left"):
        fs, *_ = orc.face_interpolate(a, orc.physics(), U)
    Ue, _ = face_test_state(a.face_centroid[:, :2], 2)
    assert np.abs(fs[:, 2] - Ue[2]).max() > 1e-3


# ------------------------------------------------------------------------------------------- brute force, tiny meshes
@pytest.mark.parametrize("faces", [[2, 2], [3, 3, 3]])
def test_average_mass_flux_rhs_on_tiny_periodic_boxes(faces):
    """AverageFlux's mass flux is 1/2 (rho_L u_n,L + rho_R u_n,R) A (averageFlux.cpp:2-10), so on a closed periodic
    cell the cell's own part cancels (sum_f n_f A_f = 0) and R_rho,c = -(1 / 2V) sum_f rho_n u_n . n_f,out A_f over its
    neighbours, every face hand-checkable (2 x 2 and 3 x 3 x 3 boxes, wrap faces included)."""
    dim = len(faces)
    mesh = BoxMesh(faces, [0.0] * dim, [1.0] * dim, [True] * dim, ordering="natural")
    a = mesh.arrays
    rng = np.random.default_rng(9)
    NI = mesh.n_interior
    U = np.stack([_state(rng.uniform(0.8, 1.3), rng.uniform(-60, 60, dim), rng.uniform(8e4, 1.2e5)) for _ in range(NI)]).T
    R, _ = orc.rhs(a, orc.physics(flux="average"), U.copy())
    want = np.zeros(NI)
    for f, (l, r) in enumerate(a.face_cells):
        n = a.face_normal[f, :dim]  # area-weighted, left to right
        want[l] -= 0.5 * (U[2 : 2 + dim, r] @ n) / a.cell_volume[l]
        want[r] += 0.5 * (U[2 : 2 + dim, l] @ n) / a.cell_volume[r]
    scale = np.abs(U[2 : 2 + dim]).max() * np.abs(a.face_normal).max() / a.cell_volume[:NI].min()
    np.testing.assert_allclose(R[0], want, rtol=1e-9, atol=1e-13 * scale)
    if faces == [2, 2]:  # both faces of a direction reach the same neighbour: the mass fluxes cancel
        assert np.abs(R[0]).max() <= 1e-13 * scale


# ------------------------------------------------------------------------------------------- periodic wrap faces
def _wrap_vs_interior(n, dim, drop_shift=False, seed=11):
    """Second-order Navier-Stokes residual (least squares + Barth-Jespersen, AUSM, viscous and conduction terms) of a
    periodic n^dim box, and the residual of the same cells in the middle block of a periodic (3n)^dim box holding the
    state three times over per direction: the middle block's stencils (two cells deep) never reach a wrap face, so
    a wrap face must behave exactly like the interior face it stands for.  Returns the largest difference per field
    relative to that field's largest residual."""
    small = BoxMesh([n] * dim, [0.0] * dim, [1.0] * dim, [True] * dim, ordering="natural").arrays
    big = BoxMesh([3 * n] * dim, [-1.0] * dim, [2.0] * dim, [True] * dim, ordering="natural").arrays
    if drop_shift:  # planted mistake: wrap faces without the periodic image of their right cell
        small = MeshArrays(**{**small.__dict__, "face_shift": np.zeros_like(small.face_shift)})
    h = 1.0 / n
    ijk_s = np.floor(small.cell_centroid[:, :dim] / h).astype(int)
    ijk_b = np.floor((big.cell_centroid[:, :dim] + 1.0) / h).astype(int)
    rng = np.random.default_rng(seed)
    shape = (n,) * dim
    rho = 1.0 + 0.2 * rng.random(shape)
    vel = rng.normal(0.0, 30.0, (dim,) + shape)
    p = 1.0e5 * (1.0 + 0.1 * rng.random(shape))

    def state(ijk):
        at = tuple((ijk % n).T)
        r, v, pp = rho[at], vel[(slice(None),) + at], p[at]
        return np.vstack([r, pp / (G - 1.0) + 0.5 * r * np.sum(v * v, axis=0), r * v])

    phys = orc.physics(flux="ausm", recon="leastsquares", limiter="barthjespersen", mu=1.8e-5, k=0.026)
    Rs, _ = orc.rhs(small, phys, state(ijk_s))
    Rb, _ = orc.rhs(big, phys, state(ijk_b))
    key = lambda ijk, m: np.ravel_multi_index(tuple(ijk.T), (m,) * dim)
    where = np.empty((3 * n) ** dim, dtype=np.int64)
    where[key(ijk_b, 3 * n)] = np.arange(big.n_cells)
    Rm = Rb[:, where[key(ijk_s + n, 3 * n)]]  # the middle-block cell at the same place as each small-box cell
    scale = np.max(np.abs(Rm), axis=1)
    return np.max(np.abs(Rs - Rm), axis=1) / scale


@pytest.mark.parametrize("dim,n", [(2, 6), (3, 5)])
def test_periodic_wrap_faces_match_interior_faces(dim, n):
    """Periodic images: the geometry of a face across the wrap (face centroid and right-cell centroid in the left
    cell's frame) must make the wrap cells' gradients, limiters, face states and viscous fluxes those of an interior
    cell, as DMPlex's localized periodic coordinates give them."""
    dev = _wrap_vs_interior(n, dim)
    assert np.all(dev < 1e-10), dev


def test_periodic_wrap_planted_mistake_fails():
    dev = _wrap_vs_interior(6, 2, drop_shift=True)
    assert np.max(dev) > 1e-2, dev
