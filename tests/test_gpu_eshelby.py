"""The reference's tests/Eshelby regression (3D, five levels) on the device:
the hierarchy from the device's tags (alamo_el_tag_bbox, blocking factor 8)
equals the oracle's, the composite FP64 solve (alamo_amr_solve) with the
device-built models and Newton right-hand sides matches the oracle's
composite solve, and the checker's line-out errors against the reference
curves stay below its 0.1 tolerance (tests/Eshelby/test:6,43-92).  The
FP32-cycle composite solve reaches the same target; its difference to the
FP64 solution is stated."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import eshelby_deck as E  # noqa: E402
from alamo_amd import elastic as el  # noqa: E402
from alamo_amd import model as M  # noqa: E402
from alamo_amd import problems as P  # noqa: E402


def device_eshelby(precision="fp64"):
    base = el.ElasticOperator(E.N_CELL, E.LENGTH / E.N_CELL, M.ISOTROPIC)

    def eta_of(op):
        return torch.from_numpy(E.eta_on(op.origin, op.level_shape(0), op.dx[0])).cuda()

    ops = el.amr_levels(base, eta_of, levels=E.MAX_LEVEL + 1, threshold=E.THRESHOLD, buffer=E.BUFFER,
                        blocking_factor=E.BLOCKING_FACTOR)
    Ct, Ft = P.eshelby_models()
    Cs, Fs, bs = [], [], []
    for op in ops:
        eta = eta_of(op)
        C = op.mix_models(eta, Ct)
        F0 = op.mix_models(eta, Ft)
        op.SetModel(C)
        b = op.bc_init(torch.zeros((3,) + op.level_shape(0), dtype=torch.float64, device="cuda"))
        bs.append(op.newton_rhs(C, torch.zeros_like(b), b, F0=F0))
        Cs.append(C)
        Fs.append(F0)
    xs = [torch.zeros_like(b) for b in bs]
    info = el.amr_solve(ops, xs, bs, 1e-8, 0.0, max_iter=100, precision=precision)
    sig = [op.stress(C, x, F0=F) for op, C, x, F in zip(ops, Cs, xs, Fs)]
    torch.cuda.synchronize()
    return ops, xs, sig, info


def test_eshelby_deck_on_the_device():
    from test_eshelby_oracle import oracle_eshelby
    ops, xs, sig, info = device_eshelby()
    boxes = [(op.clo, op.chi) for op in ops[1:]]
    oboxes, geo, xo, so, io = oracle_eshelby()
    print(f"device boxes {boxes}")
    assert boxes == oboxes
    assert info.status == 0 and io["status"] == 0
    print(f"device {info.iters} iterations, oracle {io['iters']}")
    assert info.iters == io["iters"]
    for l in range(len(ops)):
        g = xs[l].cpu().numpy()
        e = np.linalg.norm(g - xo[l]) / np.linalg.norm(xo[l])
        print(f"level {l}: {ops[l].level_shape(0)} nodes, rel L2 device vs oracle {e:.2e}")
        assert e < 1e-9
    z, u = E.line_out(geo, [x.cpu().numpy() for x in xs])
    _, s = E.line_out(geo, [g[[0, 4, 8]].cpu().numpy() for g in sig])
    uerr, serr = E.checker_errors(z, u, s)
    print(f"Eshelby device: u errors {np.round(uerr, 4)}, sigma errors {np.round(serr, 4)}")
    assert max(uerr) < E.TOLERANCE and max(serr) < E.TOLERANCE


def test_eshelby_fp32_cycles_reach_the_fp64_target():
    ops, x64, _, i64 = device_eshelby("fp64")
    ops32, x32, _, i32 = device_eshelby("fp32")
    assert i64.status == 0 and i32.status == 0
    num = sum(float(torch.linalg.norm(a - b)) ** 2 for a, b in zip(x32, x64)) ** 0.5
    den = sum(float(torch.linalg.norm(b)) ** 2 for b in x64) ** 0.5
    print(f"Eshelby FP32 cycles: {i32.iters} iterations (FP64 {i64.iters}); rel L2 to FP64 {num / den:.2e}")
    assert num / den < 1e-6
