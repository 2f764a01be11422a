"""GPU parity of the operator pieces against the FP64 CPU oracle, through the
C-ABI (alamo_amd.elastic wraps include/alamo_elastic.h).

Inputs are seeded random fields on small, non-cubic boxes so that a swapped
axis or stride shows.  Tolerance: FP64 kernels must agree with the oracle to
1e-12 relative to the field's max-norm (the two differ only in summation
order)."""
import numpy as np
import pytest

from alamo_amd import bc as B
from alamo_amd import model as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

REL = 1e-12


def _el():
    from alamo_amd import elastic
    return elastic


def rand_C(sym, shape, rng):
    nc = M.NCOMP[sym]
    if sym == M.ISOTROPIC:
        C = np.stack([rng.uniform(0.5, 1.5, shape), rng.uniform(0.5, 1.5, shape)])
    elif sym == M.MAJORMINOR:
        base = M.isotropic_slots(1.0, 1.0)
        C = base[:, None, None, None] + 0.2 * rng.standard_normal((nc,) + shape)
    else:
        C = rng.uniform(0.5, 1.5, (nc,) + shape)
    return np.ascontiguousarray(C)


def rand_bc(rng, with_neumann=True):
    bc = B.Constant()
    choices = [B.DISPLACEMENT, B.TRACTION] + ([B.NEUMANN] if with_neumann else [])
    bc.type[:] = rng.choice(choices, size=(26, 3))
    bc.val[:] = rng.standard_normal((26, 3))
    return bc


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = [(sym, n_cell, uniform) for sym in (M.ISOTROPIC, M.MAJORMINOR, M.DIAGONAL)
         for n_cell, uniform in (((8, 12, 16), False), ((16, 8, 4), True))]


@pytest.mark.parametrize("sym,n_cell,uniform", CASES)
def test_fapply_diag_residual_match_oracle(sym, n_cell, uniform):
    el = _el()
    rng = np.random.default_rng(100 + sym)
    shape = (n_cell[2] + 1, n_cell[1] + 1, n_cell[0] + 1)
    dx = (1.0 / n_cell[0], 0.7 / n_cell[1], 1.3 / n_cell[2])
    C = rand_C(sym, shape, rng)
    bc = rand_bc(rng)
    u = rng.standard_normal((3,) + shape)
    b = rng.standard_normal((3,) + shape)
    op = el.ElasticOperator(n_cell, dx, sym)
    op.SetBC(bc)
    op.SetUniform(uniform)
    op.SetModel(dev(C))
    f = host(op.Fapply(dev(u)))
    assert relerr(f, O.fapply(u, C, sym, bc, dx, uniform=uniform)) < REL
    d = host(op.Diagonal())
    assert relerr(d, O.diagonal(C, sym, bc, dx, shape)) < REL
    r = host(op.residual(dev(u), dev(b)))
    assert relerr(r, O.residual(u, b, C, sym, bc, dx, uniform=uniform)) < REL


@pytest.mark.parametrize("sym", [M.ISOTROPIC, M.MAJORMINOR, M.DIAGONAL])
def test_fsmooth_matches_oracle(sym):
    el = _el()
    rng = np.random.default_rng(7 + sym)
    n_cell = (8, 8, 12)
    shape = (13, 9, 9)
    dx = 1.0 / 8
    C = rand_C(sym, shape, rng)
    bc = B.Constant()  # displacement rows: diagonal 1
    x = rng.standard_normal((3,) + shape)
    b = rng.standard_normal((3,) + shape)
    op = el.ElasticOperator(n_cell, dx, sym)
    op.SetBC(bc)
    op.SetModel(dev(C))
    op.prepare()
    xd = dev(x)
    op.Fsmooth(xd, dev(b), nsmooth=3)
    diag = O.diagonal(C, sym, bc, dx, shape)
    ref = O.fsmooth(x, b, C, sym, bc, dx, diag, nsmooth=3)
    assert relerr(host(xd), ref) < REL


def grain_C(shape, rng, ngrains=12):
    """Piecewise-constant MajorMinor field: nearest of ngrains random seeds,
    each grain its own rotated cubic stiffness (gradient mask set only near
    grain boundaries)."""
    nz, ny, nx = shape
    seeds = rng.random((ngrains, 3)) * np.array([nx, ny, nz])
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d = [(x - sx) ** 2 + (y - sy) ** 2 + (z - sz) ** 2 for sx, sy, sz in seeds]
    gid = np.argmin(np.stack(d), axis=0)
    R = M.haar_rotations(ngrains, int(rng.integers(1 << 30)))
    table = np.stack([M.cubic(1.7, 1.2, 0.75, R[g]) for g in range(ngrains)])
    return M.field_from_table(gid, table)


# Levels with at least 32 interior nodes along every axis run the LDS-tiled
# MajorMinor kernel (tiles of 32 x 8 columns, chunks of planes in k): a box
# whose interior is 35 x 33 x 39 has partial tiles along i, j and k.
@pytest.mark.parametrize("field", ["random", "grains", "uniform"])
def test_tiled_level_matches_oracle(field):
    _tiled_case(field, {"random": 11, "grains": 12, "uniform": 13}[field])


# Levels too thin for tiles run the slab kernel with the compact grad-C
# matrices in slab order: a 129 x 257 x 9-node box (298,809 nodes, many
# waves per slab) checks that path on a large level.
def test_large_slab_level_compact_gradc_matches_oracle():
    el = _el()
    rng = np.random.default_rng(17)
    n_cell = (128, 256, 8)
    shape = (9, 257, 129)
    dx = (1.0 / 128, 2.0 / 256, 0.0625 / 8)
    C = grain_C(shape, rng, ngrains=20)
    bc = rand_bc(rng)
    u = rng.standard_normal((3,) + shape)
    b = rng.standard_normal((3,) + shape)
    op = el.ElasticOperator(n_cell, dx, M.MAJORMINOR)
    op.SetBC(bc)
    op.SetModel(dev(C))
    assert relerr(host(op.Fapply(dev(u))), O.fapply(u, C, M.MAJORMINOR, bc, dx)) < REL
    op.prepare()
    xd = dev(u)
    op.Fsmooth(xd, dev(b), nsmooth=1)
    diag = O.diagonal(C, M.MAJORMINOR, bc, dx, shape)
    assert relerr(host(xd), O.fsmooth(u, b, C, M.MAJORMINOR, bc, dx, diag, nsmooth=1)) < REL


# The k chunks are sized for ALAMO_EL_TILE_TARGET tile workgroups (read when
# C is set; by default about two rounds of resident workgroups, a different
# chunk length in FP64 and FP32).  The 35 x 33 x 39 interior has 10 tiles a
# plane: targets 1, 43 and 130 give chunks of 16 (the longest), 9 and 3
# planes, so that every position of a chunk's packed gradient masks is used.
@pytest.mark.parametrize("target", [1, 43, 130])
@pytest.mark.parametrize("field", ["random", "grains"])
def test_tiled_chunk_lengths_match_oracle(field, target, monkeypatch):
    monkeypatch.setenv("ALAMO_EL_TILE_TARGET", str(target))
    _tiled_case(field, 20 + target)


def _tiled_case(field, seed):
    el = _el()
    rng = np.random.default_rng(seed)
    n_cell = (36, 34, 40)
    shape = (41, 35, 37)
    dx = (1.0 / 36, 0.9 / 34, 1.1 / 40)
    sym = M.MAJORMINOR
    C = grain_C(shape, rng) if field != "random" else rand_C(sym, shape, rng)
    uniform = field == "uniform"
    if uniform:
        C = np.ascontiguousarray(np.broadcast_to(C[:, :1, :1, :1], C.shape))
    bc = rand_bc(rng)
    u = rng.standard_normal((3,) + shape)
    b = rng.standard_normal((3,) + shape)
    op = el.ElasticOperator(n_cell, dx, sym)
    op.SetBC(bc)
    op.SetUniform(uniform)
    op.SetModel(dev(C))
    # the grain and uniform fields take the palette (one row per distinct tensor), the random one cannot
    assert op.palette_size() == _palette_rows(C)
    assert (op.palette_size() > 0) == (field != "random")
    f = host(op.Fapply(dev(u)))
    assert relerr(f, O.fapply(u, C, sym, bc, dx, uniform=uniform)) < REL
    r = host(op.residual(dev(u), dev(b)))
    assert relerr(r, O.residual(u, b, C, sym, bc, dx, uniform=uniform)) < REL
    op.prepare()
    xd = dev(u)
    op.Fsmooth(xd, dev(b), nsmooth=2)
    diag = O.diagonal(C, sym, bc, dx, shape)
    assert relerr(host(xd), O.fsmooth(u, b, C, sym, bc, dx, diag, nsmooth=2, uniform=uniform)) < REL


def _palette_rows(C):
    """Distinct stiffness tensors among the nodes; 0 when the level keeps its
    C copy (more than 65536 of them, or more than one per 16 nodes)."""
    nodes = int(np.prod(C.shape[1:]))
    n = len(np.unique(np.ascontiguousarray(C.reshape(C.shape[0], -1).T), axis=0))
    return n if n <= min(65536, nodes // 16) else 0


def test_palette_gives_the_bits_of_the_c_copy(monkeypatch):
    """The tiled kernel reading C through the palette (one byte a node and a
    table of the distinct tensors) against the same kernel reading the
    tile-blocked C copy (ALAMO_EL_PALETTE=0): Fapply, residual and Fsmooth in
    FP64, and a solve with FP32 cycles, bit for bit, on a 41^3-node
    polycrystal of 16 grains."""
    el = _el()
    from alamo_amd import problems as P
    p = P.polycrystal(40, ngrains=16, seed=5)
    rng = np.random.default_rng(9)
    shape = (41, 41, 41)
    u = dev(rng.standard_normal((3,) + shape))
    b = dev(rng.standard_normal((3,) + shape))
    out = {}
    for mode in ("palette", "copy"):
        if mode == "copy":
            monkeypatch.setenv("ALAMO_EL_PALETTE", "0")
        op = el.ElasticOperator(p["n_cell"], p["dx"], p["sym"])
        op.SetBC(p["bc"])
        op.SetModel(dev(p["C"]))
        assert op.palette_size() == (_palette_rows(p["C"]) if mode == "palette" else 0) and _palette_rows(p["C"]) > 1
        op.prepare()
        xs = u.clone()
        op.Fsmooth(xs, b, nsmooth=2)
        x = torch.zeros_like(b)
        el.Linear(op, precision="fp32", cycle="V", max_iter=4, abort_on_fail=False).solve(x, b, 1e-12, 0.0)
        out[mode] = [host(op.Fapply(u)), host(op.residual(u, b)), host(xs), host(x)]
    for a, c in zip(out["palette"], out["copy"]):
        assert np.array_equal(a, c)


def test_transfers_and_coefficient_averaging_match_oracle():
    el = _el()
    rng = np.random.default_rng(3)
    n_cell = (16, 8, 12)
    shape = (13, 9, 17)
    op = el.ElasticOperator(n_cell, 1.0 / 16, M.MAJORMINOR)
    assert op.num_levels == 3  # 16x8x12 -> 8x4x6 -> 4x2x3 stops: 3 cells is odd
    fine = rng.standard_normal((3,) + shape)
    c = host(op.restriction(dev(fine), 1))
    assert relerr(c, O.restrict(fine)) < REL
    crse = rng.standard_normal((3,) + op.level_shape(1))
    fd = dev(fine)
    op.interpolation(fd, dev(crse), 0)
    assert relerr(host(fd), O.interp_add(crse, fine)) < REL
    # coefficients on MG level 1 come from averageDownCoeffsSameAmrLevel: check
    # through the level-1 diagonal
    C = rand_C(M.MAJORMINOR, shape, rng)
    op.SetModel(dev(C))
    op.prepare()
    C1 = O.average_down_C(C, M.MAJORMINOR)
    d1 = host(op.Diagonal(1))
    assert relerr(d1, O.diagonal(C1, M.MAJORMINOR, B.Constant(), 2.0 / 16, op.level_shape(1))) < REL


def test_aos_layout_equals_soa():
    el = _el()
    rng = np.random.default_rng(11)
    shape = (9, 9, 9)
    C = rand_C(M.MAJORMINOR, shape, rng)
    u = rng.standard_normal((3,) + shape)
    a = el.ElasticOperator(8, 0.125, M.MAJORMINOR)
    a.SetModel(dev(C))
    b = el.ElasticOperator(8, 0.125, M.MAJORMINOR)
    b.SetModel(dev(np.moveaxis(C, 0, -1)), layout=el.AOS)
    assert np.array_equal(host(a.Fapply(dev(u))), host(b.Fapply(dev(u))))


def test_bc_init_and_newton_rhs_match_oracle():
    el = _el()
    rng = np.random.default_rng(5)
    p = __import__("alamo_amd.problems", fromlist=["x"]).inclusion(16)
    shape = (17, 17, 17)
    bc = rand_bc(rng)
    op = el.ElasticOperator(16, p["dx"], p["sym"])
    op.SetBC(bc)
    b = host(op.bc_init(dev(np.zeros((3,) + shape))))
    assert np.array_equal(b, O.bc_init(shape, bc))
    bh = host(op.bc_init(dev(np.zeros((3,) + shape)), homogeneous=True))
    assert np.array_equal(bh, O.bc_init(shape, bc, homogeneous=True))
    u = 1e-3 * rng.standard_normal((3,) + shape)
    r = host(op.newton_rhs(dev(p["C"]), dev(u), dev(b), F0=dev(p["F0"])))
    assert relerr(r, O.newton_rhs(u, b, p["C"], p["sym"], bc, p["dx"], F0=p["F0"])) < REL


@pytest.mark.parametrize("with_u", [False, True])
def test_newton_rhs_closed_form_on_device(with_u):
    """k_newton_rhs against the closed form of the oracle's pin: C and F0
    linear in space, u quadratic; the interior rows are b - div(C:(grad u - F0))
    exactly (Newton.H:216-231, central divergence)."""
    from test_oracle_closed_forms import newton_closed_form_case
    el = _el()
    u, b, C, F0, bc, dx, want, div, inner = newton_closed_form_case(with_u)
    nz, ny, nx = u.shape[1:]
    op = el.ElasticOperator((nx - 1, ny - 1, nz - 1), dx, M.MAJORMINOR)
    op.SetBC(bc)
    r = host(op.newton_rhs(dev(C), dev(u), dev(b), F0=dev(F0)))
    err = np.abs(r[inner] - want[inner]).max() / np.abs(div[inner]).max()
    assert err < 1e-11, err
    assert relerr(r, O.newton_rhs(u, b, C, M.MAJORMINOR, bc, dx, F0=F0)) < REL


def test_shape_mismatch_is_rejected_before_launch():
    el = _el()
    op = el.ElasticOperator(8, 0.125, M.ISOTROPIC)
    with pytest.raises(el.AlamoError):
        op.Fapply(torch.zeros((3, 9, 9, 8), dtype=torch.float64, device="cuda"))
    with pytest.raises(el.AlamoError):
        op.SetModel(torch.zeros((2, 9, 9, 8), dtype=torch.float64, device="cuda"))


def test_model_mixing_matches_oracle_bitwise():
    """alamo_el_mix_models (src/Integrator/Mechanics.H:249-253) against the
    oracle: MajorMinor C slots and F0 of 3 models under a diffuse random eta,
    the inclusion problem's sharp two-material C, and the size limit."""
    el = _el()
    rng = np.random.default_rng(31)
    n_cell = (12, 10, 8)
    shape = (9, 11, 13)
    op = el.ElasticOperator(n_cell, 1.0 / 12, M.MAJORMINOR)
    eta = rng.random((3,) + shape)
    for ncomp in (21, 9):
        table = rng.standard_normal((3, ncomp))
        out = host(op.mix_models(dev(eta), table))
        assert np.array_equal(out, O.mix_models(eta, table))
    from alamo_amd import problems as P
    p = P.inclusion(16)
    ind = M.sphere_indicator(16, (0.5, 0.5, 0.5), 0.2)
    lm, mm = M.lame_from_E_nu(1.0, 0.3)
    li, mi = M.lame_from_E_nu(10.0, 0.3)
    iso = el.ElasticOperator(16, 1.0 / 16, M.ISOTROPIC)
    C = host(iso.mix_models(dev(np.stack([1.0 - ind, ind])), [[lm, mm], [li, mi]]))
    assert np.array_equal(C, p["C"])
    with pytest.raises(el.AlamoError):
        op.mix_models(dev(rng.random((13,) + shape)), rng.standard_normal((13, 21)))


@pytest.mark.parametrize("sym", [M.ISOTROPIC, M.MAJORMINOR, M.DIAGONAL])
def test_stress_strain_and_face_traction_match_oracle(sym):
    """alamo_el_stress and alamo_el_face_traction
    (src/Integrator/Base/Mechanics.H:210-245,391-448) against the oracle on a
    random field, with an eigenstrain, on all six faces."""
    el = _el()
    rng = np.random.default_rng(40 + sym)
    n_cell = (10, 8, 6)
    shape = (7, 9, 11)
    dx = (1.0 / 10, 0.8 / 8, 0.9 / 6)
    C = rand_C(sym, shape, rng)
    F0 = 1e-3 * rng.standard_normal((9,) + shape)
    u = rng.standard_normal((3,) + shape)
    op = el.ElasticOperator(n_cell, dx, sym)
    sig, eps = op.stress(dev(C), dev(u), F0=dev(F0), strain=True)
    ref = O.stress(u, C, sym, dx, F0=F0)
    assert relerr(host(sig), ref) < REL
    # grad u from the oracle: the Diagonal contraction with unit C is the identity
    gu = O.stress(u, np.ones((9,) + shape), M.DIAGONAL, dx).reshape((3, 3) + shape)
    assert relerr(host(eps), (0.5 * (gu + np.swapaxes(gu, 0, 1))).reshape((9,) + shape)) < REL
    s_host = host(sig)
    for axis in range(3):
        for side in (0, 1):
            tr, dp = op.face_traction(sig, dev(u), axis, side)
            rt, rd = O.face_traction(s_host, u, dx, axis, side)
            assert np.allclose(tr, rt, rtol=1e-12, atol=1e-14 * np.abs(rt).max())
            assert np.array_equal(dp, rd)


@pytest.mark.parametrize("sym", [M.ISOTROPIC, M.MAJORMINOR, M.DIAGONAL])
def test_psi_weighted_operator_matches_oracle(sym):
    """SetPsi (src/Operator/Elastic.cpp:118-259, 340-363, 839-907): Fapply,
    Diagonal, the residual and Fsmooth with a random cell weight match the
    oracle on the finest level and, after prepare, on MG level 1 (psi and C
    averaged down), with a random boundary table of displacement, traction and
    Neumann rows; FP64, within the operator bar."""
    el = _el()
    rng = np.random.default_rng(40 + sym)
    n_cell = (8, 12, 16)
    shape = (n_cell[2] + 1, n_cell[1] + 1, n_cell[0] + 1)
    dx = (1.0 / 8, 0.7 / 12, 1.3 / 16)
    C = rand_C(sym, shape, rng)
    bc = rand_bc(rng)
    psi = rng.uniform(0.05, 1.0, (shape[0] - 1, shape[1] - 1, shape[2] - 1))
    small = 1e-3
    op = el.ElasticOperator(n_cell, dx, sym)
    op.SetBC(bc)
    op.SetModel(dev(C))
    op.SetPsi(dev(psi), small)
    op.prepare()
    levels = [(C, psi, shape, dx)]
    Cc = O.average_down_C(C, sym)
    levels.append((Cc, O.psi_restrict(psi), Cc.shape[1:], tuple(2 * h for h in dx)))
    for mglev, (Cl, pl, sh, h) in enumerate(levels):
        u = rng.standard_normal((3,) + sh)
        b = rng.standard_normal((3,) + sh)
        f = host(op.Fapply(dev(u), mglev=mglev))
        assert relerr(f, O.fapply_psi(u, Cl, sym, bc, h, pl, small)) < REL
        d = host(op.Diagonal(mglev=mglev))
        assert relerr(d, O.diagonal_psi(Cl, sym, bc, h, sh, pl, small)) < REL
        r = host(op.residual(dev(u), dev(b), mglev=mglev))
        assert relerr(r, O.residual_psi(u, b, Cl, sym, bc, h, pl, small)) < REL
        xd = dev(u)
        op.Fsmooth(xd, dev(b), mglev=mglev, nsmooth=2)
        assert relerr(host(xd), O.fsmooth_psi(u, b, Cl, sym, bc, h, pl, small, nsmooth=2)) < REL


def test_psi_on_a_tiled_palette_level_matches_oracle():
    """The psi weight on a level served by the LDS-tiled kernel through the C
    palette (a 37 x 35 x 41-node grain field): Fapply, the residual and
    Fsmooth match the oracle's psi restatement."""
    el = _el()
    rng = np.random.default_rng(31)
    n_cell = (36, 34, 40)
    shape = (41, 35, 37)
    dx = (1.0 / 36, 0.9 / 34, 1.1 / 40)
    sym = M.MAJORMINOR
    C = grain_C(shape, rng)
    bc = rand_bc(rng)
    psi = rng.uniform(0.1, 1.0, (shape[0] - 1, shape[1] - 1, shape[2] - 1))
    u = rng.standard_normal((3,) + shape)
    b = rng.standard_normal((3,) + shape)
    op = el.ElasticOperator(n_cell, dx, sym)
    op.SetBC(bc)
    op.SetModel(dev(C))
    op.SetPsi(dev(psi), 1e-4)
    assert op.palette_size() == _palette_rows(C) > 0
    assert relerr(host(op.Fapply(dev(u))), O.fapply_psi(u, C, sym, bc, dx, psi, 1e-4)) < REL
    assert relerr(host(op.residual(dev(u), dev(b))), O.residual_psi(u, b, C, sym, bc, dx, psi, 1e-4)) < REL
    op.prepare()
    xd = dev(u)
    op.Fsmooth(xd, dev(b), nsmooth=1)
    assert relerr(host(xd), O.fsmooth_psi(u, b, C, sym, bc, dx, psi, 1e-4, nsmooth=1)) < REL


def test_multibox_field_gathered_into_the_single_box_layout():
    """A field held as several nodal boxes (an amrex::MultiFab's FABs on one
    level: neighbouring boxes share their face nodes) is gathered into the
    level's single-box layout by alamo_el_box_copy and scattered back, bit for
    bit; the solve of the gathered right-hand side is the solve of the
    original."""
    el = _el()
    from alamo_amd import problems as P
    rng = np.random.default_rng(23)
    p = P.polycrystal(32, ngrains=8, seed=2)
    op = el.ElasticOperator(p["n_cell"], p["dx"], p["sym"])
    op.SetBC(p["bc"])
    op.SetModel(dev(p["C"]))
    shape = op.level_shape(0)
    full = dev(rng.standard_normal((3,) + shape))
    # four FABs: x split at node 16, y at node 12 (cells 0..15 | 16..31, 0..11 | 12..31)
    boxes = [((0, 0, 0), (16, 12, 32)), ((16, 0, 0), (32, 12, 32)), ((0, 12, 0), (16, 32, 32)),
             ((16, 12, 0), (32, 32, 32))]
    fabs = []
    for lo, hi in boxes:
        fab = torch.empty((3, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1), dtype=torch.float64,
                          device="cuda")
        el.box_copy(fab, lo, full, (0, 0, 0), lo, hi)                  # scatter
        fabs.append(fab)
    gathered = torch.full_like(full, float("nan"))
    for (lo, hi), fab in zip(boxes, fabs):
        el.box_copy(gathered, (0, 0, 0), fab, lo, lo, hi)              # gather
    assert torch.equal(gathered, full)
    for (lo, hi), fab in zip(boxes, fabs):
        assert torch.equal(fab, full[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])
    b = op.bc_init(torch.zeros_like(full))
    x1, x2 = torch.zeros_like(b), torch.zeros_like(b)
    bg = torch.full_like(b, float("nan"))
    for lo, hi in boxes:
        fab = torch.empty((3, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1), dtype=torch.float64,
                          device="cuda")
        el.box_copy(fab, lo, b, (0, 0, 0), lo, hi)
        el.box_copy(bg, (0, 0, 0), fab, lo, lo, hi)
    el.Linear(op).solve(x1, b, 1e-10, 0.0)
    el.Linear(op).solve(x2, bg, 1e-10, 0.0)
    assert torch.equal(x1, x2)
    with pytest.raises(el.AlamoError):
        el.box_copy(gathered, (0, 0, 0), fabs[0], (0, 0, 0), (0, 0, 0), (17, 12, 32))   # beyond the FAB
