"""GPU solve parity: the plain FP64 multigrid on the GPU against the oracle's
converged FP64 solve of the same problem, plus the FP32-refinement and FGMRES
variants judged by the FP64 residual and the distance to the FP64 solution.

Bars (stated per test): plain FP64 GPU solution within 1e-10 relative L2 of
the oracle's; iteration counts equal (same algorithm, same inputs).

Damped Jacobi (omega = 2/3) on this strong-form operator is unstable on the
inclusion interface (spectral radius about 1.02 per Fsmooth), so the
reference F-cycle on the inclusion stalls near 1e-9 and then diverges, while
the V-cycle reaches 1e-10; the tolerances below are the ones each cycle
reaches on each problem."""
import numpy as np
import pytest

from alamo_amd import bc as B
from alamo_amd import model as M
from alamo_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def gpu_setup(p):
    from alamo_amd import elastic as el
    op = el.ElasticOperator(p["n_cell"], p["dx"], p["sym"])
    op.SetBC(p["bc"])
    op.SetModel(dev(p["C"]))
    shape = op.level_shape(0)
    b = op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda"))
    u0 = torch.zeros_like(b)
    F0 = dev(p["F0"]) if p["F0"] is not None else None
    rhs = op.newton_rhs(dev(p["C"]), u0, b, F0=F0)
    return el, op, rhs


def oracle_rhs(p):
    shape = (p["n_cell"] + 1,) * 3
    b = O.bc_init(shape, p["bc"])
    return O.newton_rhs(np.zeros((3,) + shape), b, p["C"], p["sym"], p["bc"], p["dx"], F0=p["F0"])


# F-cycle on the inclusion: the unstable interface mode grows about 2.75x per
# cycle, so it amplifies the last-bit rounding differences between two FP64
# implementations (summation order) to ~1e-9 by the time 1e-8 is reached; the
# bar for that case is 1e-8 relative L2 and equal iteration counts.
@pytest.mark.parametrize("cycle,tol,bar", [("F", 1e-8, 1e-8), ("V", 1e-10, 1e-10)])
def test_inclusion_fp64_matches_oracle(cycle, tol, bar):
    p = P.inclusion(32)
    el, op, rhs = gpu_setup(p)
    rhs_o = oracle_rhs(p)
    assert np.abs(host(rhs) - rhs_o).max() <= 1e-12 * np.abs(rhs_o).max()
    x = torch.zeros_like(rhs)
    solver = el.Linear(op, cycle=cycle)
    solver.solve(x, rhs, tol, 0.0)
    xo, info = O.mlmg_solve(rhs_o, p["C"], p["sym"], p["bc"], p["dx"], tol_rel=tol, cycle=cycle)
    assert info["status"] == 0 and solver.info.status == 0
    assert solver.info.iters == info["iters"]
    assert rel_l2(host(x), xo) < bar


# 16 cells: every level runs the per-node slab kernel; 40 cells: the finest
# level (39^3 interior nodes) runs the LDS-tiled kernel.
@pytest.mark.parametrize("n_cell", [16, 40])
def test_polycrystal_fp64_matches_oracle(n_cell):
    p = P.polycrystal(n_cell, ngrains=8, seed=3)
    el, op, rhs = gpu_setup(p)
    rhs_o = oracle_rhs(p)
    x = torch.zeros_like(rhs)
    solver = el.Linear(op)
    solver.solve(x, rhs, 1e-8, 0.0)
    xo, info = O.mlmg_solve(rhs_o, p["C"], p["sym"], p["bc"], p["dx"], tol_rel=1e-8)
    assert solver.info.iters == info["iters"]
    assert rel_l2(host(x), xo) < 1e-10


# The direct bottom solves exactly, with the FP64 inverse of the operator of
# the finest MG level with at most 4096 unknowns (9^3 nodes here), inverted on
# the device by Gauss-Jordan.  Three FP64 V-cycles against the oracle's MLMG
# coarsened to the same level (max_coarsening_level) with its BiCGStab bottom
# run to 1e-14 on those 2187 unknowns (a direct solve to rounding): the two
# agree to 1e-10.
# ALAMO_EL_DIRECT_MAX=375 moves the direct bottom to the 5^3 level (375 unknowns).
# The inverse comes from block Gauss-Jordan (pivoting inside 64-row blocks,
# checked against the operator) or, with ALAMO_EL_INVERSE=pivoted, from the
# panels pivoted over all rows.
@pytest.mark.parametrize("n_cell,coarsenings,direct_max,inverse", [(16, 1, None, "blocked"), (32, 2, None, "blocked"),
                                                                    (16, 2, "375", "blocked"),
                                                                    (16, 1, None, "pivoted"), (16, 2, "375", "pivoted")])
def test_direct_bottom_cycles_match_oracle(n_cell, coarsenings, direct_max, inverse, monkeypatch):
    if direct_max:
        monkeypatch.setenv("ALAMO_EL_DIRECT_MAX", direct_max)
    if inverse == "pivoted":
        monkeypatch.setenv("ALAMO_EL_INVERSE", "pivoted")
    p = P.polycrystal(n_cell, ngrains=8, seed=4)
    el, op, rhs = gpu_setup(p)
    rhs_o = oracle_rhs(p)
    x = torch.zeros_like(rhs)
    s = el.Linear(op, cycle="V", bottom_solver="direct", max_iter=3, abort_on_fail=False)
    s.solve(x, rhs, 1e-14, 0.0)
    xo, info = O.mlmg_solve(rhs_o, p["C"], p["sym"], p["bc"], p["dx"], tol_rel=1e-14, max_iter=3, cycle="V",
                            bottom_tol_rel=1e-14, bottom_max_iter=4000, max_coarsening_level=coarsenings)
    assert s.info.iters == 3 and info["iters"] == 3
    assert rel_l2(host(x), xo) < 1e-10


# Direct levels beyond the cubic hierarchies, for the other shapes of the
# device inversion: 9 x 9 x 11 nodes (2673 unknowns; pivoted: 8-column panels
# of 6 rows a thread) and 9 x 9 x 13 (3159; a partial last block of 23 rows;
# pivoted: 8 rows a thread).  Three FP64
# V-cycles with the direct bottom against the same cycles with the
# hierarchy stopped at that level and BiCGStab run to 1e-14 there (a direct
# solve to rounding), both on the device: they agree to 1e-10.
@pytest.mark.parametrize("inverse", ["blocked", "pivoted"])
@pytest.mark.parametrize("nz", [20, 24])
def test_direct_bottom_on_non_cubic_levels(nz, inverse, monkeypatch):
    if inverse == "pivoted":
        monkeypatch.setenv("ALAMO_EL_INVERSE", "pivoted")
    from alamo_amd import elastic as el
    shape = (nz + 1, 17, 17)
    z, y, xg = np.meshgrid(*[np.linspace(0.0, 1.0, s) for s in shape], indexing="ij")
    E = 1.0 + 0.5 * np.sin(3.0 * xg + 2.0 * y) * np.cos(2.0 * z)
    lam, mu = M.lame_from_E_nu(1.0, 0.3)
    C = dev(np.ascontiguousarray(np.stack([lam * E, mu * E])))
    bc = P.polycrystal(16, ngrains=2, seed=0)["bc"]
    xs = []
    for levels, bottom in ((2, "direct"), (1, "bicgstab")):
        op = el.ElasticOperator((16, 16, nz), 1.0 / 16, M.ISOTROPIC, max_coarsening_level=levels)
        op.SetBC(bc)
        op.SetModel(C)
        b = op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda"))
        rhs = op.newton_rhs(C, torch.zeros_like(b), b)
        x = torch.zeros_like(rhs)
        s = el.Linear(op, cycle="V", bottom_solver=bottom, max_iter=3, abort_on_fail=False, bottom_tol_rel=1e-14,
                      bottom_max_iter=4000)
        s.solve(x, rhs, 1e-14, 0.0)
        assert s.info.iters == 3
        xs.append(host(x))
    assert rel_l2(xs[0], xs[1]) < 1e-10


@pytest.mark.parametrize("precision,outer,n_cell", [("fp32", "mlmg", 32), ("fp64", "fgmres", 32),
                                                    ("fp32", "fgmres", 32), ("fp32", "fgmres", 40)])
def test_variants_reach_the_fp64_residual(precision, outer, n_cell):
    p = P.polycrystal(n_cell, ngrains=16, seed=1)
    el, op, rhs = gpu_setup(p)
    x64 = torch.zeros_like(rhs)
    el.Linear(op, outer="fgmres").solve(x64, rhs, 1e-11, 0.0)
    x = torch.zeros_like(rhs)
    s = el.Linear(op, precision=precision, outer=outer)
    s.solve(x, rhs, 1e-8, 0.0)
    target = 1e-8 * max(s.info.rhsnorm, s.info.res0)
    r = op.residual(x, rhs)
    assert op.norm_inf(r) <= target * (1 + 1e-9)
    # distance to the FP64 solution, stated: the solve stops at 1e-8 relative
    # residual, so it agrees with the 1e-11 solution to about the conditioning
    # of A times that
    assert rel_l2(host(x), host(x64)) < 1e-5


def test_uniaxial_tension_closed_form():
    """tests/UniaxialTension/test:27-28 restated on one level: u_x = eps x,
    u_y = -nu eps y, u_z = -nu eps z reproduce exactly."""
    from alamo_amd import elastic as el
    n = 16
    lam, mu = M.lame_from_E_nu(210.0, 0.3)
    shape = (n + 1,) * 3
    C = np.stack([np.full(shape, lam), np.full(shape, mu)])
    eps = 0.1 / 16
    bc = B.uniaxial(0, eps)
    op = el.ElasticOperator(n, 1.0 / n, M.ISOTROPIC)
    op.SetBC(bc)
    op.SetModel(dev(C))
    b = op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda"))
    x = torch.zeros_like(b)
    el.Linear(op).solve(x, b, 1e-10, 0.0)
    z, y, xx = M.node_coords(n)
    ue = np.stack([np.broadcast_to(eps * xx, shape), np.broadcast_to(-0.3 * eps * y, shape),
                   np.broadcast_to(-0.3 * eps * z, shape)])
    assert np.abs(host(x) - ue).max() < 1e-8 * eps


@pytest.mark.parametrize("name,F0", [("linear", None), ("affine", 1e-3)])
def test_solid_cubic_regression_end_to_end(name, F0):
    """tests/Solid (input:7-27, test:24-51) on the GPU end to end: for each
    time of the committed reference series, the device Newton right-hand side
    (src/Solver/Nonlocal/Newton.H:151-231), the FP64 solve, the device stress
    (src/Integrator/Base/Mechanics.H:210-245) and the x-hi face traction of
    thermo.dat (Mechanics.H:391-448); the checker's metric against the
    reference's trac_xhi_x must stay below its 1e-4 tolerance."""
    import os
    from alamo_amd import elastic as el
    ref = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", f"solid_{name}_cubic.dat"), skiprows=1)
    n = 4
    N = n + 1
    slots = M.cubic(168.3, 1.221, 0.757)
    C = dev(np.ascontiguousarray(np.broadcast_to(slots[:, None, None, None], (21, N, N, N))))
    F0f = None
    if F0 is not None:
        F0f = dev(np.ascontiguousarray(np.broadcast_to((F0 * np.eye(3)).ravel()[:, None, None, None], (9, N, N, N))))
    op = el.ElasticOperator(n, 1.0 / n, M.MAJORMINOR)
    op.SetModel(C)
    zero = torch.zeros((3, N, N, N), dtype=torch.float64, device="cuda")
    trac = []
    for t in ref[:, 0]:
        op.SetBC(B.tension_test("uniaxial_stress", disp=0.01 * t))
        b = op.bc_init(zero.clone())
        rhs = op.newton_rhs(C, zero, b, F0=F0f)
        du = torch.zeros_like(b)
        el.Linear(op, cycle="V").solve(du, rhs, 1e-10, 0.0)
        sig = op.stress(C, du, F0=F0f)
        trac.append(op.face_traction(sig, du, axis=0, side=1)[0][0])
    trac = np.array(trac)

    def integrate(xv, yv):
        return np.sum(0.5 * (yv[1:] + yv[:-1]) * (xv[1:] - xv[:-1]))
    times = ref[:, 0]
    err = np.sqrt(integrate(times, (ref[:, 2] - trac) ** 2)) / np.sqrt(integrate(times, ref[:, 2] ** 2))
    assert err < 1e-4, err


@pytest.mark.parametrize("precision,outer", [("fp64", "mlmg"), ("fp32", "mlmg"), ("fp32", "fgmres")])
def test_psi_weighted_solve_solves_the_weighted_system(precision, outer):
    """SetPsi on a 33^3-node polycrystal: a diffuse sphere of psi = 1 in
    psi = 0.2 (tanh edge 0.1 wide), psi_small 1e-8.  The V-cycle solve (FP64,
    FP32 cycles in FP64 refinement, FGMRES) converges, and the oracle's
    weighted residual of the device solution, computed on the CPU, meets the
    solver's own target.  A level whose coarsest MG level holds more unknowns
    than the weighted dense bottom is rejected before any launch."""
    from alamo_amd import elastic as el
    p = P.polycrystal(32, ngrains=8, seed=2)
    _, op, rhs = gpu_setup(p)
    n = p["n_cell"]
    zc, yc, xc = np.meshgrid(*[(np.arange(n) + 0.5) / n] * 3, indexing="ij")
    r = np.sqrt((xc - 0.5) ** 2 + (yc - 0.5) ** 2 + (zc - 0.5) ** 2)
    psi = 0.2 + 0.8 * 0.5 * (1 - np.tanh((r - 0.35) / 0.05))
    op.SetPsi(dev(psi), 1e-8)
    x = torch.zeros_like(rhs)
    s = el.Linear(op, precision=precision, outer=outer, cycle="V")
    s.solve(x, rhs, 1e-9, 0.0)
    print(f"{precision} {outer}: {s.info.iters} iterations, |r| {s.info.res:.3e} of |b| {s.info.rhsnorm:.3e}")
    assert s.info.status == 0
    res = O.residual_psi(host(x), host(rhs), p["C"], p["sym"], p["bc"], p["dx"], psi, 1e-8)
    assert np.abs(res).max() <= 1e-9 * max(s.info.rhsnorm, s.info.res0) * (1 + 1e-6)
    # 18^3 cells coarsen once, to 10^3 nodes = 3000 unknowns: too many for the dense bottom
    big = el.ElasticOperator(18, 1.0 / 18, M.ISOTROPIC)
    big.SetModel(torch.ones((2, 19, 19, 19), dtype=torch.float64, device="cuda"))
    big.SetPsi(torch.ones((18, 18, 18), dtype=torch.float64, device="cuda"))
    with pytest.raises(el.AlamoError):
        el.Linear(big).solve(torch.zeros((3, 19, 19, 19), dtype=torch.float64, device="cuda"),
                             torch.ones((3, 19, 19, 19), dtype=torch.float64, device="cuda"), 1e-8, 0.0)


def test_prepare_fp32_ahead_of_the_solve_gives_the_same_bits():
    """alamo_el_prepare_fp32 makes the FP32 copies the first FP32 solve would
    make; a solve after it is the same, bit for bit, after a model update too."""
    p = P.polycrystal(32, ngrains=8, seed=2)
    el, op, rhs = gpu_setup(p)
    C = dev(p["C"])
    xs = []
    for ahead in (False, True, True):
        op.SetModel(C)
        op.prepare()
        if ahead:
            op.prepare_fp32()
        x = torch.zeros_like(rhs)
        el.Linear(op, precision="fp32", outer="fgmres", bottom_solver="direct").solve(x, rhs, 1e-8, 0.0)
        xs.append(x)
    assert torch.equal(xs[0], xs[1]) and torch.equal(xs[1], xs[2])


@pytest.mark.parametrize("outer", ["mlmg", "fgmres"])
def test_max_fmg_iter_and_fixed_iter_match_oracle(outer):
    """Linear::max_fmg_iter (F-cycles for the first k iterations, V-cycles
    after) and fixed_iter (at most k iterations, no failure;
    src/Solver/Nonlocal/Linear.H:266-267): the FP64 MLMG solve takes the
    oracle's iteration count and solution (1e-10); the FGMRES solve with the
    switching preconditioner reaches the target; fixed_iter ends a solve whose
    target is out of reach after k iterations, at the oracle's residual."""
    p = P.polycrystal(16, ngrains=8, seed=4)
    el, op, rhs = gpu_setup(p)
    rhs_o = oracle_rhs(p)
    x = torch.zeros_like(rhs)
    s = el.Linear(op, outer=outer, max_fmg_iter=2)
    s.solve(x, rhs, 1e-10, 0.0)
    xo, info = O.mlmg_solve(rhs_o, p["C"], p["sym"], p["bc"], p["dx"], tol_rel=1e-10, max_fmg_iter=2)
    if outer == "mlmg":
        assert s.info.iters == info["iters"]
        assert rel_l2(host(x), xo) < 1e-10
    else:
        assert s.info.status == 0
        assert op.norm_inf(op.residual(x, rhs)) <= 1e-10 * max(s.info.rhsnorm, s.info.res0) * (1 + 1e-9)
    if outer == "mlmg":
        x = torch.zeros_like(rhs)
        s = el.Linear(op, fixed_iter=3)  # abort_on_fail: a failure would raise
        s.solve(x, rhs, 1e-300, 0.0)
        _, info3 = O.mlmg_solve(rhs_o, p["C"], p["sym"], p["bc"], p["dx"], tol_rel=1e-300, fixed_iter=3)
        assert s.info.status == 0 and s.info.iters == 3 and info3["iters"] == 3
        assert abs(s.info.res - info3["res"]) <= 1e-6 * info3["res"]


@pytest.mark.parametrize("precision,restart,n_cell", [("fp64", 10, 32), ("fp32", 10, 40), ("fp32", 3, 32),
                                                      ("fp32", 12, 40)])
def test_delayed_reorthogonalization_matches_cgs2(precision, restart, n_cell, monkeypatch):
    """FGMRES with restart <= 12 delays each basis vector's second
    Gram-Schmidt projection to the next step (DCGS2: two passes over the basis
    a step instead of three).  Against the undelayed passes (ALAMO_EL_GS=cgs2)
    on the same problem: both reach the target, the step counts differ by at
    most one, and both solutions lie within 1e-5 relative L2 of an FP64 solve
    to 1e-11 (the solves stop at 1e-8 relative residual).  Restart 3 runs
    several restart cycles (the last column of each is kept as projected once);
    12 fills the kernels' vector limit.  The basis is FP64 in both
    (ALAMO_EL_V=fp64; the FP32 basis has a test of its own)."""
    monkeypatch.setenv("ALAMO_EL_V", "fp64")
    p = P.polycrystal(n_cell, ngrains=16, seed=3)
    el, op, rhs = gpu_setup(p)
    x64 = torch.zeros_like(rhs)
    el.Linear(op, outer="fgmres").solve(x64, rhs, 1e-11, 0.0)
    kw = dict(el.FAST_PATH, precision=precision, restart=restart, max_fmg_iter=2)
    got = {}
    for gs in ("cgs2", "dcgs2"):
        if gs == "cgs2":
            monkeypatch.setenv("ALAMO_EL_GS", "cgs2")
        else:
            monkeypatch.delenv("ALAMO_EL_GS", raising=False)
        x = torch.zeros_like(rhs)
        s = el.Linear(op, max_iter=400, **kw)
        s.solve(x, rhs, 1e-8, 0.0)
        target = 1e-8 * max(s.info.rhsnorm, s.info.res0)
        assert s.info.status == 0
        assert op.norm_inf(op.residual(x, rhs)) <= target * (1 + 1e-9)
        assert rel_l2(host(x), host(x64)) < 1e-5
        got[gs] = s.info.iters
    assert abs(got["cgs2"] - got["dcgs2"]) <= 1, got


@pytest.mark.parametrize("n_cell,cycle", [(32, "F"), (40, "F"), (40, "V")])
def test_fp32_z_basis_gives_the_bits_of_the_fp64_copy(n_cell, cycle, monkeypatch):
    """The FP32 path keeps its preconditioned directions Z in FP32 (the FP32
    cycle's output, whose widening to FP64 is exact): the FP64 Fapply reads
    them directly (tiled levels, 40^3 here) or through an FP64 copy (slab
    levels, 32^3), a V-cycle writes its correction straight into Z, and the
    solution update widens them.  The solve is the one with Z stored in FP64
    (ALAMO_EL_Z=fp64), bit for bit: same iterations, same solution.  (The
    basis V stays FP64 in both, ALAMO_EL_V=fp64: an FP64 Z keeps it FP64.)"""
    from alamo_amd import elastic as el
    monkeypatch.setenv("ALAMO_EL_V", "fp64")
    p = P.polycrystal(n_cell, ngrains=16, seed=5)
    _, op, rhs = gpu_setup(p)
    kw = dict(el.FAST_PATH, cycle=cycle, max_fmg_iter=2)
    xs, its = [], []
    for z in ("fp64", "fp32"):
        if z == "fp64":
            monkeypatch.setenv("ALAMO_EL_Z", "fp64")
        else:
            monkeypatch.delenv("ALAMO_EL_Z", raising=False)
        x = torch.zeros_like(rhs)
        s = el.Linear(op, max_iter=400, **kw)
        s.solve(x, rhs, 1e-8, 0.0)
        assert s.info.status == 0
        xs.append(x)
        its.append(s.info.iters)
    assert its[0] == its[1]
    assert torch.equal(xs[0], xs[1])


@pytest.mark.parametrize("n_cell,cycle,restart", [(32, "F", 10), (40, "F", 10), (40, "V", 4), (40, "F", 3)])
def test_steps_launched_ahead_give_the_bits_of_the_synchronous_loop(n_cell, cycle, restart, monkeypatch):
    """With the delayed passes the next FGMRES step is launched before the
    host has read the last one's coefficients (its scale w / |w| is read on
    the device).  A step launched ahead is kept, or discarded when the cycle
    ends (restarts of 3 and 4 end cycles often): the solve is the
    synchronous loop's (ALAMO_EL_AHEAD=0), bit for bit, in FP64 and with the
    FP32 cycles."""
    from alamo_amd import elastic as el
    p = P.polycrystal(n_cell, ngrains=16, seed=7)
    _, op, rhs = gpu_setup(p)
    for prec in ("fp64", "fp32"):
        kw = dict(el.FAST_PATH, cycle=cycle, max_fmg_iter=2, restart=restart)
        if prec == "fp64":
            kw["precision"] = "fp64"
        xs, its = [], []
        for ahead in ("0", "1"):
            monkeypatch.setenv("ALAMO_EL_AHEAD", ahead)
            x = torch.zeros_like(rhs)
            s = el.Linear(op, max_iter=400, **kw)
            s.solve(x, rhs, 1e-8, 0.0)
            assert s.info.status == 0
            xs.append(x)
            its.append(s.info.iters)
        print(f"{n_cell}^3 {cycle} restart {restart} {prec}: {its[0]} steps")
        assert its[0] == its[1]
        assert torch.equal(xs[0], xs[1])


@pytest.mark.parametrize("n_cell,cycle,restart,steps", [(32, "F", 10, 1), (40, "F", 10, 1), (40, "V", 10, 1),
                                                         (40, "F", 3, 1), (40, "F", 12, 2)])
def test_fp32_krylov_basis_reaches_the_fp64_target(n_cell, cycle, restart, steps, monkeypatch):
    """With the FP32 cycles and the delayed passes the Krylov basis V is kept
    in FP32 (sums in FP64).  Against the FP64 basis (ALAMO_EL_V=fp64) on the
    same problem: both meet the FP64 residual target, both solutions lie
    within 1e-6 relative L2 of an FP64 solve to 1e-11, and the step counts
    differ by at most one (two at restart 12, whose longer cycles lose more
    to the FP32 rounding of the basis).  Restart 3 runs many restart cycles,
    each starting from the true FP64 residual."""
    from alamo_amd import elastic as el
    p = P.polycrystal(n_cell, ngrains=16, seed=9)
    _, op, rhs = gpu_setup(p)
    x64 = torch.zeros_like(rhs)
    el.Linear(op, outer="fgmres").solve(x64, rhs, 1e-11, 0.0)
    kw = dict(el.FAST_PATH, cycle=cycle, max_fmg_iter=2, restart=restart)
    got = {}
    for v in ("fp64", "fp32"):
        if v == "fp64":
            monkeypatch.setenv("ALAMO_EL_V", "fp64")
        else:
            monkeypatch.delenv("ALAMO_EL_V", raising=False)
        x = torch.zeros_like(rhs)
        s = el.Linear(op, max_iter=400, **kw)
        s.solve(x, rhs, 1e-8, 0.0)
        target = 1e-8 * max(s.info.rhsnorm, s.info.res0)
        assert s.info.status == 0
        assert op.norm_inf(op.residual(x, rhs)) <= target * (1 + 1e-9)
        e = rel_l2(host(x), host(x64))
        print(f"{n_cell}^3 {cycle} restart {restart} V {v}: {s.info.iters} steps, rel L2 to FP64 {e:.2e}")
        assert e < 1e-6
        got[v] = s.info.iters
    assert abs(got["fp64"] - got["fp32"]) <= steps, got
