"""GPU tests of the learned coarse-level solver inside the elastic solve.

- the level-L right-hand side and the network input against host
  restatements built from the oracle's restriction;
- the committed trained network (alamo_amd/weights, trained on the GPU by
  `python -m alamo_amd.train` on 1024 solver-generated samples) on held-out
  129^3 polycrystals: what it gains, i.e. the error of its initial guess
  against the FP64 solution (1.0 for the zero guess), and the iterations and
  time of the FP32-refined solve with and without it;
- the judge: every learned mode reaches the FP64 solution, and an untrained
  network cannot make the solve take more cycles than it does without it.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from alamo_amd import bc as B  # noqa: E402
from alamo_amd import elastic as el  # noqa: E402
from alamo_amd import model as M  # noqa: E402
from alamo_amd import problems  # noqa: E402
from oracle import oracle as O  # noqa: E402

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "alamo_amd", "weights",
                       "unet_polycrystal129_level2.npz")


def _op(p):
    op = el.ElasticOperator(p["n_cell"], p["dx"], p["sym"])
    op.SetBC(p["bc"])
    op.SetModel(torch.from_numpy(p["C"]).cuda())
    op.prepare()
    return op


def test_coarse_rhs_and_features_match_host_restatement():
    p = problems.polycrystal(32, ngrains=8, seed=3)
    p["bc"] = B.shear(2, 0, 1e-3)
    op = _op(p)
    rng = np.random.default_rng(0)
    b = op.bc_init(torch.from_numpy(rng.standard_normal((3, 33, 33, 33)) * 0.01).cuda())
    L = 1
    bc = op.coarse_rhs(b, L)
    bh = b.cpu().numpy()
    bint = bh.copy()
    bint[:, [0, -1]] = 0
    bint[:, :, [0, -1]] = 0
    bint[:, :, :, [0, -1]] = 0
    ref = O.restrict(bint)
    bnd = np.zeros(ref.shape, dtype=bool)
    bnd[:, [0, -1]] = bnd[:, :, [0, -1]] = bnd[:, :, :, [0, -1]] = True
    ref[bnd] = bh[:, ::2, ::2, ::2][bnd]
    assert np.allclose(bc.cpu().numpy(), ref, rtol=1e-14, atol=1e-18)

    feat, s = op.learned_input(bc, L)
    C = op.level_model(L).cpu().numpy()
    alpha = C[0].mean()
    t = bc.cpu().numpy().copy()
    t[:, ~bnd[0]] /= alpha
    assert np.isclose(s, np.abs(t).max(), rtol=1e-14)
    f = feat.cpu().numpy()
    nc = 21
    assert np.allclose(f[:nc], C / alpha, rtol=1e-6)
    assert np.allclose(f[nc:nc + 3], np.where(bnd, 0, t / s), rtol=1e-6, atol=1e-7)
    assert np.allclose(f[nc + 3:nc + 6], np.where(bnd, t / s, 0), rtol=1e-6, atol=1e-7)
    # displacement rows: the shear table clamps the z faces, other faces are traction free
    zface = np.zeros(bnd.shape[1:], dtype=bool)
    zface[[0, -1]] = True
    assert np.array_equal(f[nc + 6] == 1.0, zface)
    assert np.allclose(f[nc + 9], np.linspace(0, 1, 17)[None, None, :])
    faces = [t[:, :, :, 0], t[:, :, 0, :], t[:, 0, :, :], t[:, :, :, -1], t[:, :, -1, :], t[:, -1, :, :]]
    means = np.array([[fc[c].mean() for c in range(3)] for fc in faces]).ravel() / s
    assert np.allclose(f[nc + 12:nc + 30, 0, 0, 0], means, rtol=1e-6, atol=1e-8)


@pytest.fixture(scope="module")
def trained():
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained weights")
    from alamo_amd import train
    net, meta = train.load(WEIGHTS)
    return net, meta


@pytest.mark.parametrize("seed", [7, 11])
def test_trained_network_guess_on_held_out_polycrystal(trained, seed):
    net, meta = trained
    p = problems.polycrystal(128, ngrains=64, seed=seed)
    op = _op(p)
    op.set_learned(net, meta["level"])
    b = op.bc_init(torch.zeros((3,) + op.level_shape(0), dtype=torch.float64, device="cuda"))
    x64 = torch.zeros_like(b)
    el.Linear(op, cycle="F", outer="fgmres").solve(x64, b, 1e-10, 0.0)
    g = op.learned_guess(b)
    gam = float(torch.sum(g * x64) / torch.sum(g * g))
    err = float(torch.linalg.norm(gam * g - x64) / torch.linalg.norm(x64))
    print(f"seed {seed}: best-scaled learned guess rel L2 error {err:.4f} (zero guess: 1.0)")
    assert err < 0.2

    import time
    res = {}
    for mode in ("", "guess", "guess+precond"):
        x = torch.zeros_like(b)
        s = el.Linear(op, cycle="F", precision="fp32", outer="fgmres", bottom_solver="direct", learned=mode)
        torch.cuda.synchronize()
        t = time.perf_counter()
        s.solve(x, b, 1e-8, 0.0)
        torch.cuda.synchronize()
        res[mode] = (s.info.iters, time.perf_counter() - t, s.info.learned_steps)
        e = float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))
        print(f"  learned={mode!r:16s} iters {s.info.iters} learned steps {s.info.learned_steps} "
              f"{res[mode][1] * 1e3:.1f} ms, rel L2 vs FP64 {e:.2e}")
        assert s.info.status == 0 and e < 1e-7
    # the judge keeps the learned steps from costing cycles
    assert res["guess"][0] <= res[""][0] + 1 and res["guess+precond"][0] <= res[""][0] + 1


def test_untrained_network_cannot_slow_the_solve():
    from alamo_amd.learned import UNet
    p = problems.polycrystal(64, ngrains=16, seed=5)
    op = _op(p)
    net = UNet((17, 17, 17), 21 + 30, 3, seed=1)
    op.set_learned(net, 2)
    b = op.bc_init(torch.zeros((3,) + op.level_shape(0), dtype=torch.float64, device="cuda"))
    its = {}
    for mode in ("", "guess+precond"):
        for cyc, outer in (("F", "fgmres"), ("V", "mlmg")):
            x = torch.zeros_like(b)
            s = el.Linear(op, cycle=cyc, outer=outer, learned=mode)
            s.solve(x, b, 1e-8, 0.0)
            its[(mode, outer)] = s.info.iters
            assert s.info.status == 0
    print(its)
    assert its[("guess+precond", "fgmres")] <= its[("", "fgmres")] + 1
    assert its[("guess+precond", "mlmg")] <= its[("", "mlmg")] + 1


def test_set_learned_rejects_mismatched_network():
    from alamo_amd.learned import UNet
    p = problems.polycrystal(32, ngrains=4, seed=1)
    op = _op(p)
    with pytest.raises(el.AlamoError):
        op.set_learned(UNet((17, 17, 17), 21 + 30, 3, seed=0), 0)  # level 0 has 33^3 nodes
    with pytest.raises(el.AlamoError):
        op.set_learned(UNet((17, 17, 17), 21 + 29, 3, seed=0), 1)
    iso = el.ElasticOperator(32, 1 / 32, M.ISOTROPIC)
    iso.SetBC(B.Constant())
    with pytest.raises(el.AlamoError):
        iso.set_learned(UNet((17, 17, 17), 21 + 30, 3, seed=0), 1)
