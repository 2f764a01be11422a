"""The red-black Gauss-Seidel smoother (ALAMO_EL_SMOOTH_RBGS; no counterpart
in the reference, whose smoother is damped Jacobi).

- alamo_el_gs_smooth against a numpy restatement over the oracle's Fapply and
  Diagonal: per half sweep, the nodes of one color (i + j + k) % 2 take
  x += w (b - A x) / D with w = 0.9 inside and 1/2 on boundary rows, the others
  keep their values; 1e-12 relative, three symmetries, random boundary types.
- Its only parity gate beyond that is the converged solution: the FP64 MLMG
  with it reaches the same solution as with the reference's Jacobi (1e-8
  relative L2 at 1e-10 residual), for V and F cycles, and the FP32-cycle
  FGMRES path with it reaches the FP64 residual target.
"""
import numpy as np
import pytest

from alamo_amd import bc as B
from alamo_amd import model as M
from alamo_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def rbgs_restated(x, b, C, sym, bc, dx, nsweep):
    shape = x.shape[1:]
    k, j, i = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    color = (i + j + k) & 1
    bnd = (i == 0) | (j == 0) | (k == 0) | (i == shape[2] - 1) | (j == shape[1] - 1) | (k == shape[0] - 1)
    w = np.where(bnd, 0.5, 0.9)
    D = O.diagonal(C, sym, bc, dx, shape)
    for _ in range(nsweep):
        for c in (0, 1):
            r = b - O.fapply(x, C, sym, bc, dx)
            x = x + np.where(color == c, w * r / D, 0.0)
    return x


@pytest.mark.parametrize("sym", [M.ISOTROPIC, M.MAJORMINOR, M.DIAGONAL])
def test_gs_smooth_matches_restatement(sym):
    from alamo_amd import elastic as el
    rng = np.random.default_rng(40 + sym)
    n_cell = (12, 10, 14)
    shape = (15, 11, 13)
    dx = (0.1, 0.12, 0.08)
    nc = M.NCOMP[sym]
    if sym == M.MAJORMINOR:
        C = M.cubic(1.7, 1.2, 0.75)[:, None, None, None] * rng.uniform(0.8, 1.2, shape)[None]
    elif sym == M.ISOTROPIC:
        C = rng.uniform(0.5, 1.5, (nc,) + shape)
    else:
        C = rng.uniform(0.5, 1.5, (nc,) + shape)
    C = np.ascontiguousarray(C)
    bc = B.Constant()
    bc.type[:] = rng.choice([B.DISPLACEMENT, B.TRACTION, B.NEUMANN], size=(26, 3))
    op = el.ElasticOperator(n_cell, dx, sym)
    op.SetBC(bc)
    op.SetModel(torch.from_numpy(C).cuda())
    x0 = rng.standard_normal((3,) + shape)
    b = rng.standard_normal((3,) + shape)
    xd = torch.from_numpy(x0.copy()).cuda()
    op.gs_smooth(xd, torch.from_numpy(b).cuda(), nsmooth=2)
    torch.cuda.synchronize()
    want = rbgs_restated(x0, b, C, sym, bc, dx, 2)
    err = np.abs(xd.cpu().numpy() - want).max() / np.abs(want).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("name,cycle", [("polycrystal", "F"), ("polycrystal", "V"), ("inclusion", "V")])
def test_rbgs_cycles_reach_the_jacobi_solution(name, cycle):
    from alamo_amd import elastic as el
    p = P.polycrystal(32, ngrains=16, seed=1) if name == "polycrystal" else P.inclusion(32)
    op = el.ElasticOperator(p["n_cell"], p["dx"], p["sym"])
    op.SetBC(p["bc"])
    C = torch.from_numpy(p["C"]).cuda()
    op.SetModel(C)
    b = op.bc_init(torch.zeros((3,) + op.level_shape(0), dtype=torch.float64, device="cuda"))
    F0 = torch.from_numpy(p["F0"]).cuda() if p["F0"] is not None else None
    rhs = op.newton_rhs(C, torch.zeros_like(b), b, F0=F0)
    xs = {}
    for sm in ("jacobi", "rbgs"):
        x = torch.zeros_like(rhs)
        s = el.Linear(op, cycle=cycle, smoother=sm, max_iter=300)
        s.solve(x, rhs, 1e-10, 0.0)
        print(f"{name} {cycle}-cycle, {sm}: {s.info.iters} iterations")
        assert s.info.status == 0
        xs[sm] = x
    e = float(torch.linalg.norm(xs["rbgs"] - xs["jacobi"]) / torch.linalg.norm(xs["jacobi"]))
    print(f"rel L2 red-black GS vs Jacobi solution {e:.2e}")
    assert e < 1e-8
    x = torch.zeros_like(rhs)
    s = el.Linear(op, cycle="F", precision="fp32", outer="fgmres", smoother="rbgs")
    s.solve(x, rhs, 1e-8, 0.0)
    assert s.info.status == 0
    assert op.norm_inf(op.residual(x, rhs)) <= 1e-8 * max(s.info.rhsnorm, s.info.res0) * (1 + 1e-9)
