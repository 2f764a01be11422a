"""Pins of the FP64 CPU oracle (oracle/elastic_oracle.cpp) that do not retype
its formulas: brute-force tensor sums from the reference's own slot table,
closed forms that the discretisation reproduces exactly, assembled-matrix
identities, and the reference regression data of tests/Solid.
"""
import json
import os

import numpy as np
import pytest

from alamo_amd import bc as B
from alamo_amd import model as M
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def uid_full(slots):
    """3x3x3x3 tensor from 21 slots through the reference accessor table
    (tests/golden/majorminor_uid.json, extracted from
    src/Set/Matrix4_MajorMinor.H:376-423)."""
    table = json.load(open(os.path.join(GOLDEN, "majorminor_uid.json")))
    T = np.empty((3, 3, 3, 3))
    for e in table:
        i, j, k, l = e["ijkl"]
        T[i, j, k, l] = slots[e["slot"]]
    return T


def grid(n):
    return M.node_coords(n)


def test_slot_convention_matches_reference_table():
    table = json.load(open(os.path.join(GOLDEN, "majorminor_uid.json")))
    for e in table:
        assert M.mm_slot(*e["ijkl"]) == e["slot"]


def quadratic_field(n, rng):
    """u_p = sum of random quadratic monomials: central 2nd and mixed
    differences are exact on it."""
    z, y, x = grid(n)
    N = n + 1
    coef = rng.standard_normal((3, 10))
    mons = [np.ones((N, N, N)), x + 0 * y + 0 * z, y + 0 * x + 0 * z, z + 0 * x + 0 * y, x * x + 0 * y + 0 * z,
            y * y + 0 * x + 0 * z, z * z + 0 * x + 0 * y, x * y + 0 * z, x * z + 0 * y, y * z + 0 * x]
    mons = [np.broadcast_to(m, (N, N, N)) for m in mons]
    u = np.stack([sum(coef[p, q] * mons[q] for q in range(10)) for p in range(3)])
    # exact Hessian H[p][a][b] (constant)
    H = np.zeros((3, 3, 3))
    for p in range(3):
        H[p, 0, 0], H[p, 1, 1], H[p, 2, 2] = 2 * coef[p, 4], 2 * coef[p, 5], 2 * coef[p, 6]
        H[p, 0, 1] = H[p, 1, 0] = coef[p, 7]
        H[p, 0, 2] = H[p, 2, 0] = coef[p, 8]
        H[p, 1, 2] = H[p, 2, 1] = coef[p, 9]
    return u, H


@pytest.mark.parametrize("sym", [M.ISOTROPIC, M.MAJORMINOR, M.DIAGONAL])
def test_fapply_interior_is_exact_contraction_on_quadratics(sym):
    """f_i = C_ijkl u_{k,lj} for uniform C and quadratic u, against an 81-term
    brute-force sum over the full tensor."""
    rng = np.random.default_rng(sym)
    n = 8
    N = n + 1
    u, H = quadratic_field(n, rng)
    if sym == M.MAJORMINOR:
        slots = rng.uniform(0.5, 1.5, 21)
        Cfull = uid_full(slots)
        Cn = slots
    elif sym == M.ISOTROPIC:
        lam, mu = 0.7, 1.3
        Cfull = uid_full(M.isotropic_slots(lam, mu))
        Cn = np.array([lam, mu])
    else:
        A = rng.uniform(0.5, 1.5, (3, 3))
        Cfull = np.zeros((3, 3, 3, 3))
        for i in range(3):
            for j in range(3):
                Cfull[i, j, i, j] = A[i, j]
        Cn = A.ravel()
    C = np.ascontiguousarray(np.broadcast_to(Cn[:, None, None, None], (len(Cn), N, N, N)))
    f = O.fapply(u, C, sym, B.Constant(), 1.0 / n)
    expect = np.einsum("ijkl,klj->i", Cfull, H)
    inner = f[:, 1:-1, 1:-1, 1:-1]
    assert np.abs(inner - expect[:, None, None, None]).max() < 1e-9 * np.abs(expect).max()


def test_isotropic_and_majorminor_paths_agree():
    rng = np.random.default_rng(1)
    n = 6
    N = n + 1
    lam = rng.uniform(0.5, 1.5, (N, N, N))
    mu = rng.uniform(0.5, 1.5, (N, N, N))
    Ci = np.stack([lam, mu])
    Cm = np.einsum("s,xyz->sxyz", M.isotropic_slots(1.0, 0.0), lam) + np.einsum("s,xyz->sxyz",
                                                                               M.isotropic_slots(0.0, 1.0), mu)
    bc = B.Constant()
    bc.type[:] = rng.choice([B.DISPLACEMENT, B.TRACTION, B.NEUMANN], size=(26, 3))
    u = rng.standard_normal((3, N, N, N))
    a = O.fapply(u, Ci, M.ISOTROPIC, bc, 1.0 / n)
    b = O.fapply(u, np.ascontiguousarray(Cm), M.MAJORMINOR, bc, 1.0 / n)
    assert np.abs(a - b).max() < 1e-10 * np.abs(a).max()


def test_gradient_of_C_term_exact_for_linear_C():
    """f_i = C_ijkl(x) u_{k,lj} + C_ijkl,j u_{k,l}: with C linear in space and u
    quadratic both terms are exact; brute force over the full tensors."""
    rng = np.random.default_rng(2)
    n = 8
    N = n + 1
    z, y, x = grid(n)
    u, H = quadratic_field(n, rng)
    s0 = rng.uniform(1.0, 2.0, 21)
    g = rng.uniform(-0.3, 0.3, (3, 21))
    X = [np.broadcast_to(x, (N, N, N)), np.broadcast_to(y, (N, N, N)), np.broadcast_to(z, (N, N, N))]
    C = np.stack([s0[s] + g[0, s] * X[0] + g[1, s] * X[1] + g[2, s] * X[2] for s in range(21)])
    f = O.fapply(u, np.ascontiguousarray(C), M.MAJORMINOR, B.Constant(), 1.0 / n)
    h = 1.0 / n
    # grad u is linear; central differences are exact
    for (k, j, i) in [(3, 4, 5), (1, 1, 1), (7, 2, 6)]:
        xyz = np.array([i * h, j * h, k * h])
        gu = np.array([[(u[p, k, j, i + 1] - u[p, k, j, i - 1]) / (2 * h),
                        (u[p, k, j + 1, i] - u[p, k, j - 1, i]) / (2 * h),
                        (u[p, k + 1, j, i] - u[p, k - 1, j, i]) / (2 * h)] for p in range(3)])
        Cx = uid_full(s0 + g.T @ xyz)
        dC = [uid_full(g[a]) for a in range(3)]
        expect = np.einsum("ijkl,klj->i", Cx, H) + sum(np.einsum("ikl,kl->i", dC[a][:, a], gu) for a in range(3))
        assert np.allclose(f[:, k, j, i], expect, rtol=1e-10, atol=1e-10)


def test_traction_rows_are_stress_times_normal():
    """Boundary rows: f = sigma n with sigma = C : grad u, exact for affine u,
    on every region, normals with the reference's truncated constants."""
    rng = np.random.default_rng(3)
    n = 4
    N = n + 1
    z, y, x = grid(n)
    G = rng.standard_normal((3, 3))
    u = np.stack([np.broadcast_to(G[p, 0] * x + G[p, 1] * y + G[p, 2] * z, (N, N, N)) for p in range(3)])
    slots = rng.uniform(0.5, 1.5, 21)
    C = np.ascontiguousarray(np.broadcast_to(slots[:, None, None, None], (21, N, N, N)))
    bc = B.Constant()
    bc.type[:] = B.TRACTION
    f = O.fapply(u, C, M.MAJORMINOR, bc, 1.0 / n)
    sig = np.einsum("ijkl,kl->ij", uid_full(slots), G)
    S2, S3 = 0.70710678118, 0.57735026919
    checks = {(0, 2, 2): [-1, 0, 0], (n, 2, 2): [1, 0, 0], (2, 0, 2): [0, -1, 0], (2, 2, n): [0, 0, 1],
              (0, 0, 2): [-S2, -S2, 0], (2, n, n): [0, S2, S2], (n, 2, 0): [S2, 0, -S2],
              (0, 0, 0): [-S3, -S3, -S3], (n, n, n): [S3, S3, S3], (n, 0, n): [S3, -S3, S3]}
    for (i, j, k), nrm in checks.items():
        assert np.allclose(f[:, k, j, i], sig @ np.array(nrm), rtol=1e-12, atol=1e-12), (i, j, k)


def test_laplacian_model_is_textbook_laplacian():
    """Model::Solid::Linear::Laplacian (src/Model/Solid/Linear/Laplacian.H:26-58):
    C_ijkl = d_ik d_jl gives f_i = u_i,jj, the 7-point Laplacian."""
    rng = np.random.default_rng(4)
    n = 6
    N = n + 1
    u = rng.standard_normal((3, N, N, N))
    C = np.ones((9, N, N, N))
    h = 0.37
    f = O.fapply(u, C, M.DIAGONAL, B.Constant(), h)
    lap = (u[:, 1:-1, 1:-1, 2:] + u[:, 1:-1, 1:-1, :-2] + u[:, 1:-1, 2:, 1:-1] + u[:, 1:-1, :-2, 1:-1] +
           u[:, 2:, 1:-1, 1:-1] + u[:, :-2, 1:-1, 1:-1] - 6 * u[:, 1:-1, 1:-1, 1:-1]) / h ** 2
    assert np.abs(f[:, 1:-1, 1:-1, 1:-1] - lap).max() < 1e-11 * np.abs(lap).max()


@pytest.mark.parametrize("sym", [M.ISOTROPIC, M.MAJORMINOR, M.DIAGONAL])
def test_diagonal_equals_assembled_matrix_diagonal(sym):
    """Diagonal() against the diagonal of A assembled column by column (all
    nodes, interior and every boundary type)."""
    rng = np.random.default_rng(10 + sym)
    n = 3
    N = n + 1
    nc = M.NCOMP[sym]
    C = rng.uniform(0.5, 1.5, (nc, N, N, N))
    bc = B.Constant()
    bc.type[:] = rng.choice([B.DISPLACEMENT, B.TRACTION, B.NEUMANN], size=(26, 3))
    dx = (0.3, 0.5, 0.7)
    d = O.diagonal(C, sym, bc, dx, (N, N, N)).ravel()
    m = 3 * N ** 3
    for q in range(m):
        e = np.zeros(m)
        e[q] = 1.0
        col = O.fapply(e.reshape((3, N, N, N)), C, sym, bc, dx).ravel()
        assert abs(col[q] - d[q]) <= 1e-12 * max(1.0, abs(d[q])), q


def test_restriction_is_scaled_transpose_of_interpolation_inside():
    """Interior full weighting R = P^T / 8; P reproduces linear fields; R
    reproduces constants in every branch (interior, faces, edges, corners)."""
    nf = 9
    Nc = 5
    z, y, x = grid(nf - 1)
    lin = np.broadcast_to(0.3 + 1.1 * x - 0.7 * y + 0.2 * z, (nf,) * 3)
    zc, yc, xc = grid(Nc - 1)
    linc = np.broadcast_to(0.3 + 1.1 * xc - 0.7 * yc + 0.2 * zc, (Nc,) * 3)
    assert np.allclose(O.interp_add(linc[None], np.zeros((1,) + (nf,) * 3))[0], lin, atol=1e-14)
    assert np.allclose(O.restrict(np.ones((1,) + (nf,) * 3)), 1.0, atol=1e-15)
    # P^T/8 at interior coarse nodes
    P = np.zeros((nf ** 3, Nc ** 3))
    for q in range(Nc ** 3):
        e = np.zeros(Nc ** 3)
        e[q] = 1.0
        P[:, q] = O.interp_add(e.reshape((1,) + (Nc,) * 3), np.zeros((1,) + (nf,) * 3)).ravel()
    R = np.zeros((Nc ** 3, nf ** 3))
    for q in range(nf ** 3):
        e = np.zeros(nf ** 3)
        e[q] = 1.0
        R[:, q] = O.restrict(e.reshape((1,) + (nf,) * 3)).ravel()
    inner = [((K * Nc) + J) * Nc + I for K in range(1, Nc - 1) for J in range(1, Nc - 1) for I in range(1, Nc - 1)]
    assert np.allclose(R[inner], P.T[inner] / 8.0, atol=1e-15)


def test_coefficient_average_down_uses_restriction_weights():
    rng = np.random.default_rng(6)
    Cf = rng.standard_normal((21, 9, 9, 9))
    assert np.array_equal(O.average_down_C(Cf, M.MAJORMINOR), O.restrict(Cf))


def test_jacobi_fixed_point_is_the_solution():
    """The exact solution of A x = b (dense solve of the assembled matrix) is
    a fixed point of Fsmooth."""
    rng = np.random.default_rng(7)
    n = 3
    N = n + 1
    C = np.stack([rng.uniform(0.5, 1.5, (N, N, N)), rng.uniform(0.5, 1.5, (N, N, N))])
    bc = B.Constant()
    m = 3 * N ** 3
    A = np.zeros((m, m))
    for q in range(m):
        e = np.zeros(m)
        e[q] = 1.0
        A[:, q] = O.fapply(e.reshape((3, N, N, N)), C, M.ISOTROPIC, bc, 1.0 / n).ravel()
    b = rng.standard_normal(m)
    xs = np.linalg.solve(A, b).reshape((3, N, N, N))
    d = O.diagonal(C, M.ISOTROPIC, bc, 1.0 / n, (N, N, N))
    x1 = O.fsmooth(xs, b.reshape((3, N, N, N)), C, M.ISOTROPIC, bc, 1.0 / n, d, nsmooth=2)
    assert np.abs(x1 - xs).max() < 1e-12 * np.abs(xs).max()


def test_stencil_accuracy_on_cos_field():
    """src/Test/Numeric/Stencil.H:57-106 restated through Fapply: on the
    field of IC::Trig2 with n = 1 on the unit cube, cos(pi x)cos(pi y)cos(pi z)
    (src/IC/Trig2.H:89-98), at 32 cells the Laplacian-model rows and the
    isotropic rows (mixed derivatives) have inf-norm relative error < 1e-2."""
    n = 32
    N = n + 1
    z, y, x = grid(n)
    c = np.cos(np.pi * x) * np.cos(np.pi * y) * np.cos(np.pi * z)
    u = np.stack([c, np.zeros_like(c), np.zeros_like(c)])
    f = O.fapply(u, np.ones((9, N, N, N)), M.DIAGONAL, B.Constant(), 1.0 / n)
    exact = -3 * np.pi ** 2 * c
    err = np.abs(f[0, 1:-1, 1:-1, 1:-1] - exact[1:-1, 1:-1, 1:-1]).max() / np.abs(exact).max()
    assert err < 1e-2
    lam, mu = 0.6, 0.4
    f = O.fapply(u, np.stack([np.full((N, N, N), lam), np.full((N, N, N), mu)]), M.ISOTROPIC, B.Constant(), 1.0 / n)
    # f_1 = (lam + mu) u_0,01 = (lam+mu) d2c/dxdy
    dxy = (lam + mu) * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.cos(np.pi * z)
    err = np.abs(f[1, 1:-1, 1:-1, 1:-1] - dxy[1:-1, 1:-1, 1:-1]).max() / np.abs(dxy).max()
    assert err < 1e-2


def test_uniaxial_tension_closed_form():
    """tests/UniaxialTension/test:27-28: u_x = eps x, sigma_xx = E eps, lateral
    contraction -nu eps; exact for the discretisation (linear field)."""
    n = 16
    N = n + 1
    lam, mu = M.lame_from_E_nu(210.0, 0.3)
    C = np.stack([np.full((N, N, N), lam), np.full((N, N, N), mu)])
    eps = 0.1 / 16
    bc = B.uniaxial(0, eps)
    b = O.bc_init((N, N, N), bc)
    xs, info = O.mlmg_solve(b, C, M.ISOTROPIC, bc, 1.0 / n, tol_rel=1e-10)
    assert info["status"] == 0
    z, y, x = grid(n)
    ue = np.stack([np.broadcast_to(eps * x, (N, N, N)), np.broadcast_to(-0.3 * eps * y, (N, N, N)),
                   np.broadcast_to(-0.3 * eps * z, (N, N, N))])
    assert np.abs(xs - ue).max() < 1e-9 * eps
    sig = O.stress(xs, C, M.ISOTROPIC, 1.0 / n)
    assert np.allclose(sig[0], 210.0 * eps, rtol=1e-8)


def test_trig_closed_form():
    """tests/TrigTest/test:23-34: Laplacian model, body force sin sin sin,
    zero displacement; u_0 = -sin sin sin / (3 pi^2), squared relative error
    < 1e-4."""
    n = 32
    N = n + 1
    z, y, x = grid(n)
    s = np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
    b = np.zeros((3, N, N, N))
    b[0] = s
    bc = B.Constant()
    b = O.bc_init((N, N, N), bc, rhs=b)
    xs, info = O.mlmg_solve(b, np.ones((9, N, N, N)), M.DIAGONAL, bc, 1.0 / n, tol_rel=1e-10, cycle="V")
    assert info["status"] == 0
    exact = -s / (3 * np.pi ** 2)
    assert ((xs[0] - exact) ** 2).sum() / (exact ** 2).sum() < 1e-4


@pytest.mark.parametrize("name,F0", [("linear", None), ("affine", 1e-3)])
def test_solid_cubic_regression_fixture(name, F0):
    """tests/Solid (input:7-27, test:24-51) restated: 4^3 cells, cubic
    C11=168.3 C12=1.221 C44=0.757, TensionTest uniaxial_stress with x-hi
    displacement 0.01 t, trac_xhi_x integrated on the x-hi face
    (src/Integrator/Base/Mechanics.H:391-448); the checker's metric against
    the committed reference series must be below its 1e-4 tolerance."""
    ref = np.loadtxt(os.path.join(GOLDEN, f"solid_{name}_cubic.dat"), skiprows=1)
    n = 4
    N = n + 1
    slots = M.cubic(168.3, 1.221, 0.757)
    C = np.ascontiguousarray(np.broadcast_to(slots[:, None, None, None], (21, N, N, N)))
    F0f = None
    if F0 is not None:
        F0f = np.ascontiguousarray(np.broadcast_to((F0 * np.eye(3)).ravel()[:, None, None, None], (9, N, N, N)))
    h = 1.0 / n
    times = ref[:, 0]
    trac = []
    for t in times:
        bc = B.tension_test("uniaxial_stress", disp=0.01 * t)
        b = O.bc_init((N, N, N), bc)
        rhs = O.newton_rhs(np.zeros((3, N, N, N)), b, C, M.MAJORMINOR, bc, h, F0=F0f)
        du, info = O.mlmg_solve(rhs, C, M.MAJORMINOR, bc, h, tol_rel=1e-10, cycle="V")
        sig = O.stress(du, C, M.MAJORMINOR, h, F0=F0f)
        trac.append(O.face_traction(sig, du, h, axis=0, side=1)[0][0])
    trac = np.array(trac)

    def integrate(xv, yv):
        return np.sum(0.5 * (yv[1:] + yv[:-1]) * (xv[1:] - xv[:-1]))
    err = np.sqrt(integrate(times, (ref[:, 2] - trac) ** 2)) / np.sqrt(integrate(times, ref[:, 2] ** 2))
    assert err < 1e-4, err


def test_cubic_builder_pins():
    # identity rotation: C11, C12, C44 in the known slots
    s = M.cubic(1.7, 1.2, 0.75)
    assert s[M.mm_slot(0, 0, 0, 0)] == 1.7 and s[M.mm_slot(1, 1, 1, 1)] == 1.7
    assert s[M.mm_slot(0, 0, 1, 1)] == 1.2 and s[M.mm_slot(0, 1, 0, 1)] == 0.75
    # axis permutations: both conventions agree and the tensor is invariant
    R = np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], dtype=float)
    assert np.allclose(M.cubic(1.7, 1.2, 0.75, R), s)
    assert np.allclose(M.cubic(1.7, 1.2, 0.75, R, convention="tensor"), s)
    # isotropic limit of the tensor convention is rotation invariant
    Rr = M.haar_rotations(1, 3)[0]
    assert np.allclose(M.cubic(2.0, 1.0, 0.5, Rr, convention="tensor"), M.isotropic_slots(1.0, 0.5))
    # the reference convention differs from the tensor for generic rotations
    # (no C_ijji entries in its crystal-frame tensor) yet stays positive definite
    assert np.abs(M.cubic(1.7, 1.2, 0.75, Rr) - M.cubic(1.7, 1.2, 0.75, Rr, convention="tensor")).max() > 1e-2
    full = M.mm_full(M.cubic(1.7, 1.2, 0.75, Rr))
    pairs = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    w = [1, 1, 1, np.sqrt(2), np.sqrt(2), np.sqrt(2)]
    mand = np.array([[w[a] * w[b] * full[p + q] for b, q in enumerate(pairs)] for a, p in enumerate(pairs)])
    assert np.linalg.eigvalsh(mand).min() > 0


def test_haar_rotations_are_rotations():
    R = M.haar_rotations(50, 0)
    assert np.allclose(np.einsum("nij,nkj->nik", R, R), np.eye(3), atol=1e-12)
    assert np.allclose(np.linalg.det(R), 1.0)


def test_bc_parse_uniaxial_tension_deck():
    """Keys of tests/UniaxialTension/input parsed as Constant::Parse does."""
    bc = B.Constant.parse({"type.xlo": "disp trac trac", "type.xhi": "disp trac trac", "val.xhi": "0.1 0.0 0.0",
                           "type.xhiylozlo": "disp disp disp", "val.xhiylozlo": "0.1 0.0 0.0"})
    assert list(bc.type[B.FACE_INDEX["xlo"]]) == [B.DISPLACEMENT, B.TRACTION, B.TRACTION]
    assert bc.val[B.FACE_INDEX["xhi"], 0] == 0.1
    assert list(bc.type[B.FACE_INDEX["xhiylozlo"]]) == [B.DISPLACEMENT] * 3
    assert list(bc.type[B.FACE_INDEX["zlo"]]) == [B.DISPLACEMENT] * 3  # default
    with pytest.raises(KeyError):
        B.Constant.parse({"type.xlo": "disp bogus trac"})


def test_jacobi_instability_on_the_inclusion_interface():
    """Property of the reference smoother on this operator: damped Jacobi
    (omega 2/3) has spectral radius above 1 on the inclusion problem, with
    the growing mode on the inclusion interface; it is what stops the F-cycle
    from reaching 1e-10 there."""
    from alamo_amd import problems as P
    p = P.inclusion(16)
    shape = (17, 17, 17)
    d = O.diagonal(p["C"], p["sym"], p["bc"], p["dx"], shape)
    x = np.random.default_rng(0).standard_normal((3,) + shape)
    z = np.zeros_like(x)
    g = []
    for _ in range(300):
        x = O.fsmooth(x, z, p["C"], p["sym"], p["bc"], p["dx"], d)
        nrm = np.linalg.norm(x)
        g.append(nrm)
        x /= nrm
    assert np.exp(np.mean(np.log(g[-50:]))) > 1.0
    a = np.abs(x).max(axis=0)
    k, j, i = np.unravel_index(np.argmax(a), a.shape)
    r = np.sqrt(((np.array([i, j, k]) / 16.0 - 0.5) ** 2).sum())
    assert abs(r - 0.2) < 2.0 / 16


def test_model_mixing_pins():
    """Integrator::Mechanics::UpdateModel (src/Integrator/Mechanics.H:249-253)
    in the oracle: a one-hot eta returns that model's components exactly; a
    sharp two-material indicator gives the inclusion builder's C and F0 (an
    independent einsum); models that are all equal, under an eta that sums to
    one, give that model (a partition of unity, to rounding)."""
    rng = np.random.default_rng(5)
    shape = (4, 5, 6)
    table = rng.standard_normal((3, 21))
    pick = rng.integers(0, 3, shape)
    eta = np.stack([(pick == n).astype(np.float64) for n in range(3)])
    out = O.mix_models(eta, table)
    assert np.array_equal(out, np.moveaxis(table[pick], -1, 0))
    from alamo_amd import problems as P
    p = P.inclusion(8)
    ind = M.sphere_indicator(8, (0.5, 0.5, 0.5), 0.2)
    lm, mm = M.lame_from_E_nu(1.0, 0.3)
    li, mi = M.lame_from_E_nu(10.0, 0.3)
    etas = np.stack([1.0 - ind, ind])
    assert np.array_equal(O.mix_models(etas, [[lm, mm], [li, mi]]), p["C"])
    assert np.array_equal(O.mix_models(etas, [np.zeros(9), (1e-3 * np.eye(3)).ravel()]), p["F0"])
    w = rng.random((4,) + shape)
    w /= w.sum(axis=0)
    same = np.tile(table[0], (4, 1))
    assert np.allclose(O.mix_models(w, same), table[0][:, None, None, None] * np.ones(shape), rtol=1e-14, atol=1e-14)


def test_face_traction_closed_forms():
    """The oracle's thermo.dat traction (src/Integrator/Base/Mechanics.H:391-448):
    a uniform stress integrates to sigma . e_axis times the face area on every
    face; a stress linear along the face integrates exactly (the 4-corner
    average is the cell-centre value); disp_hi is u at in-face indices n - 2."""
    rng = np.random.default_rng(8)
    shape = (5, 6, 7)  # nz, ny, nx
    dx = (0.5, 0.25, 0.2)
    S = rng.standard_normal((3, 3))
    sig = np.ascontiguousarray(np.broadcast_to(S.reshape(9, 1, 1, 1), (9,) + shape))
    u = rng.standard_normal((3,) + shape)
    L = [(shape[2] - 1) * dx[0], (shape[1] - 1) * dx[1], (shape[0] - 1) * dx[2]]
    for axis in range(3):
        area = np.prod([L[a] for a in range(3) if a != axis])
        for side in (0, 1):
            tr, dp = O.face_traction(sig, u, dx, axis, side)
            assert np.allclose(tr, S[:, axis] * area, rtol=1e-13, atol=1e-13)
    z, y, x = np.meshgrid(np.arange(shape[0]) * dx[2], np.arange(shape[1]) * dx[1], np.arange(shape[2]) * dx[0],
                          indexing="ij")
    lin = np.zeros((9,) + shape)
    lin[0] = 1.0 + 2.0 * y + 3.0 * z  # sigma_xx on the x faces, linear in y and z
    tr, dp = O.face_traction(lin, u, dx, 0, 1)
    exact = L[1] * L[2] + 2.0 * L[1] ** 2 / 2 * L[2] + 3.0 * L[2] ** 2 / 2 * L[1]
    assert abs(tr[0] - exact) < 1e-12 * abs(exact) and tr[1] == 0.0 and tr[2] == 0.0
    assert np.array_equal(dp, u[:, shape[0] - 2, shape[1] - 2, shape[2] - 1])


def test_refinement_tagging_pins():
    """The oracle's TagCellsForRefinement (src/Integrator/Mechanics.H:298-313):
    a linear eta has the exact gradient everywhere (central and one-sided), so
    every cell is tagged just above the threshold and none just below; the
    nodal gradient agrees with a direct loop over the reference's stencils
    (src/Numeric/Stencil.H:70-191) on a random field; one tagged node gives
    its cell grown by the buffer and clipped."""
    shape = (5, 6, 7)  # nz, ny, nx
    dx = (0.5, 0.25, 0.2)
    dxn = np.sqrt(sum(d * d for d in dx))
    z, y, x = np.meshgrid(np.arange(shape[0]) * dx[2], np.arange(shape[1]) * dx[1], np.arange(shape[2]) * dx[0],
                          indexing="ij")
    slope = np.array([0.3, -0.4, 1.2])
    eta = (slope[0] * x + slope[1] * y + slope[2] * z)[None]
    thr = np.linalg.norm(slope) * dxn
    box, n = O.tag_box(eta, dx, threshold=thr * (1 - 1e-9), buffer=0)
    assert n == 4 * 5 * 6 and box == ((0, 0, 0), (6, 5, 4))
    assert O.tag_box(eta, dx, threshold=thr * (1 + 1e-9))[1] == 0
    rng = np.random.default_rng(4)
    f = rng.standard_normal(shape)
    n3 = (shape[2], shape[1], shape[0])
    for a, ax in enumerate((2, 1, 0)):
        g = np.gradient(f, dx[a], axis=ax, edge_order=1)
        for idx in np.ndindex(*shape):
            p = idx[ax]
            up = list(idx); dn = list(idx)
            up[ax] = min(p + 1, n3[a] - 1); dn[ax] = max(p - 1, 0)
            h = dx[a] if p in (0, n3[a] - 1) else 2 * dx[a]
            assert abs(g[idx] - (f[tuple(up)] - f[tuple(dn)]) / h) <= 1e-14 * (1 + abs(g[idx]))
    spike = np.zeros((1,) + shape)
    spike[0, 2, 3, 4] = 1.0  # the cells around node (4, 3, 2) see its gradient
    box, n = O.tag_box(spike, dx, threshold=1e-6, buffer=1)
    assert n > 0 and all(box[0][a] >= 0 and box[1][a] <= (6, 5, 4)[a] for a in range(3))
    assert box[0] == (2, 1, 0) and box[1] == (6, 5, 4)


def test_psi_weight_pins():
    """The oracle's SetPsi weight (src/Operator/Elastic.cpp:118-259, 340-363,
    839-907) pinned by mathematics:
    - psi = 1 with psi_small = 0 leaves Fapply and Diagonal unchanged, bit for bit;
    - a linear psi with uniform isotropic C and u = (xy, yz, zx) gives, at
      every interior node, the product rule div(psi' sigma) = psi' div(sigma)
      + sigma . grad psi' with psi' = (1 - s) psi + s, from closed forms;
    - a constant psi c scales the traction rows of a boundary node by
      (1 - s)(n c + 8 - n)/8 + s, n the node's cells inside the domain (the
      cells outside count as 1), and leaves the displacement rows alone;
    - the restriction keeps linear fields (value of the fine cell 2I, 2J, 2K)
      and spreads unit spikes with the reference's weights."""
    from alamo_amd import bc as B
    from alamo_amd import model as M
    n_cell = (6, 8, 10)
    shape = (n_cell[2] + 1, n_cell[1] + 1, n_cell[0] + 1)
    dx = (0.5, 0.25, 0.2)
    lam, mu = 1.3, 0.7
    C = np.empty((2,) + shape)
    C[0], C[1] = lam, mu
    z, y, x = np.meshgrid(*[np.arange(shape[2 - a]) * dx[a] for a in (2, 1, 0)], indexing="ij")
    u = np.stack([x * y, y * z, z * x])
    bc = B.Constant()
    rng = np.random.default_rng(3)
    bc.type[:] = rng.choice([B.DISPLACEMENT, B.TRACTION, B.NEUMANN], size=bc.type.shape)
    ones = np.ones((shape[0] - 1, shape[1] - 1, shape[2] - 1))
    assert np.array_equal(O.fapply_psi(u, C, M.ISOTROPIC, bc, dx, ones, small=0.0), O.fapply(u, C, M.ISOTROPIC, bc, dx))
    assert np.array_equal(O.diagonal_psi(C, M.ISOTROPIC, bc, dx, shape, ones, small=0.0),
                          O.diagonal(C, M.ISOTROPIC, bc, dx, shape))
    # product rule on a linear psi sampled at the cell centres
    g, c0, s = np.array([0.8, -0.5, 1.1]), 2.0, 1e-3
    zc, yc, xc = np.meshgrid(*[(np.arange(shape[2 - a] - 1) + 0.5) * dx[a] for a in (2, 1, 0)], indexing="ij")
    psi = c0 + g[0] * xc + g[1] * yc + g[2] * zc
    f = O.fapply_psi(u, C, M.ISOTROPIC, bc, dx, psi, small=s)
    ps = (1 - s) * (c0 + g[0] * x + g[1] * y + g[2] * z) + s
    tr = x + y + z
    sig = [[lam * tr + 2 * mu * y, mu * x, mu * z], [mu * x, lam * tr + 2 * mu * z, mu * y],
           [mu * z, mu * y, lam * tr + 2 * mu * x]]
    inner = np.s_[1:-1, 1:-1, 1:-1]
    for p in range(3):
        want = ps * (lam + mu) + (1 - s) * sum(sig[p][a] * g[a] for a in range(3))
        assert np.allclose(f[p][inner], want[inner], rtol=1e-12, atol=1e-12)
    # constant psi on the boundary: traction rows scaled, the cells outside read 1
    cst = 0.3 * ones
    f1 = O.fapply_psi(u, C, M.ISOTROPIC, bc, dx, cst, small=s)
    f0 = O.fapply(u, C, M.ISOTROPIC, bc, dx)
    d1 = O.diagonal_psi(C, M.ISOTROPIC, bc, dx, shape, cst, small=s)
    d0 = O.diagonal(C, M.ISOTROPIC, bc, dx, shape)
    reg = O.regions(shape)
    for node, n_in in (((0, 0, 0), 1), ((0, 0, 3), 2), ((0, 4, 3), 4), ((5, 4, 3), 8)):
        w = (1 - s) * (n_in * 0.3 + 8 - n_in) / 8 + s
        for p in range(3):
            scaled = reg[node] < 0 or bc.type[reg[node], p] == B.TRACTION
            if reg[node] >= 0:
                assert f1[(p,) + node] == pytest.approx(w * f0[(p,) + node] if scaled else f0[(p,) + node], rel=1e-14)
            assert d1[(p,) + node] == pytest.approx(w * d0[(p,) + node] if scaled else d0[(p,) + node], rel=1e-14)
    # restriction: linear fields are kept, spikes spread with the reference's weights
    fz, fy, fx = np.meshgrid(np.arange(8.0), np.arange(6.0), np.arange(4.0), indexing="ij")
    lin = 1.0 + 0.5 * fx - 2.0 * fy + 0.25 * fz
    assert np.allclose(O.psi_restrict(lin), lin[::2, ::2, ::2], rtol=0, atol=1e-13)
    spike = np.zeros((8, 6, 4))
    spike[2, 3, 2] = 1.0                      # (x, y, z) = (2, 3, 2): a face neighbour of coarse (1, 1, 1) and (1, 2, 1)
    r = O.psi_restrict(spike)
    assert r[1, 1, 1] == 1 / 16 and r[1, 2, 1] == 1 / 16 and np.count_nonzero(r) == 2
    spike = np.zeros((8, 6, 4))
    spike[3, 2, 0] = 1.0                      # on the low x face: coarse (0, 1, 1) and (0, 1, 2) in their face rule
    r = O.psi_restrict(spike)
    assert r[1, 1, 0] == 2 / 16 and r[2, 1, 0] == 2 / 16 and np.count_nonzero(r) == 2
    spike = np.zeros((8, 6, 4))
    spike[0, 0, 0] = 1.0                      # the corner cell: coarse (0, 0, 0) only, by copy
    r = O.psi_restrict(spike)
    assert r[0, 0, 0] == 1.0 and np.count_nonzero(r) == 1


def test_max_fmg_iter_and_fixed_iter_compose_the_cycles():
    """Linear::max_fmg_iter and fixed_iter (src/Solver/Nonlocal/Linear.H:266-267,
    305-312) on the oracle: a solve with max_fmg_iter = k is k F-cycles and then
    the V-cycle solve started from that iterate, with the same absolute target
    (the same iteration count and residual history); fixed_iter = k stops after
    k iterations without failure, on the first k steps of the unlimited
    history.  The oracle's threaded BiCGStab bottom sums in a varying order,
    so histories agree to 1e-6 relative, not bit for bit."""
    from alamo_amd import problems
    p = problems.polycrystal(16, ngrains=8, seed=4)
    shape = (17,) * 3
    b = O.bc_init(shape, p["bc"])
    rhs = O.newton_rhs(np.zeros((3,) + shape), b, p["C"], p["sym"], p["bc"], p["dx"])
    args = (rhs, p["C"], p["sym"], p["bc"], p["dx"])
    T = 1e-9
    k = 2
    x, info = O.mlmg_solve(*args, tol_rel=0.0, tol_abs=T, max_fmg_iter=k)
    xf, infof = O.mlmg_solve(*args, tol_rel=0.0, tol_abs=1e-300, cycle="F", fixed_iter=k)
    assert infof["status"] == 0 and infof["iters"] == k
    xv, infov = O.mlmg_solve(*args, x0=xf, tol_rel=0.0, tol_abs=T, cycle="V")
    assert info["status"] == 0 and infov["status"] == 0
    assert info["iters"] == k + infov["iters"]
    np.testing.assert_allclose(info["hist"], np.concatenate([infof["hist"], infov["hist"][1:]]), rtol=1e-6)
    assert np.linalg.norm(x - xv) <= 1e-6 * np.linalg.norm(x)
    # more F-cycles than needed: the same as F-cycles throughout
    xa, infoa = O.mlmg_solve(*args, tol_rel=0.0, tol_abs=T, max_fmg_iter=10**6)
    xb, infob = O.mlmg_solve(*args, tol_rel=0.0, tol_abs=T)
    assert infoa["iters"] == infob["iters"]
    assert np.linalg.norm(xa - xb) <= 1e-6 * np.linalg.norm(xb)
    # fixed_iter on a target out of reach
    xc, infoc = O.mlmg_solve(*args, tol_rel=0.0, tol_abs=1e-300, fixed_iter=3)
    assert infoc["status"] == 0 and infoc["iters"] == 3
    np.testing.assert_allclose(infoc["hist"], infob["hist"][:4], rtol=1e-6)
