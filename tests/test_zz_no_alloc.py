"""alamo_el_newton_rhs_noalloc (the Newton right-hand side through the
library's fixed per-GPU scratch, no allocation) against alamo_el_newton_rhs,
through the C-ABI: bit for bit on two calls in a row, the second on a smaller
grid; a 129^3 MajorMinor polycrystal with eigenstrain and AoS C (five slabs of
z planes), then a 33^3 isotropic inclusion with SoA C (one slab); planes
larger than the scratch are rejected."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from alamo_amd import elastic as el  # noqa: E402
from alamo_amd import model as M  # noqa: E402
from alamo_amd import problems as P  # noqa: E402


def _case(p, layout, seed):
    rng = np.random.default_rng(seed)
    op = el.ElasticOperator(p["n_cell"], p["dx"], p["sym"])
    bc = p["bc"]
    bc.type[:] = rng.choice([0, 1, 3], size=bc.type.shape)
    op.SetBC(bc)
    shape = op.level_shape(0)
    C = torch.from_numpy(p["C"]).cuda()
    if layout == el.AOS:
        C = C.permute(1, 2, 3, 0).contiguous()
    F0 = torch.from_numpy(1e-3 * rng.standard_normal((9,) + shape)).cuda()
    u = torch.from_numpy(1e-3 * rng.standard_normal((3,) + shape)).cuda()
    b = torch.from_numpy(rng.standard_normal((3,) + shape)).cuda()
    return op, C, F0, u, b


def test_noalloc_newton_rhs_matches_the_allocating_entry_point():
    cases = [(P.polycrystal(128, ngrains=64, seed=3), el.AOS, 1), (P.inclusion(32), el.SOA, 2)]
    for p, layout, seed in cases:
        op, C, F0, u, b = _case(p, layout, seed)
        ref = op.newton_rhs(C, u, b, F0=F0, layout=layout)
        out = torch.full_like(ref, float("nan"))
        op.newton_rhs(C, u, b, F0=F0, layout=layout, out=out, noalloc=True)
        torch.cuda.synchronize()
        print(f"{p['name']}: max |noalloc - newton_rhs| = {float((out - ref).abs().max()):.3e}")
        assert torch.equal(out, ref)
        # without F0 too
        ref = op.newton_rhs(C, u, b, layout=layout)
        op.newton_rhs(C, u, b, layout=layout, out=out, noalloc=True)
        assert torch.equal(out, ref)


def test_noalloc_newton_rhs_rejects_planes_beyond_the_scratch():
    op = el.ElasticOperator((399, 399, 2), 0.01, M.ISOTROPIC, max_coarsening_level=0)
    shape = op.level_shape(0)
    z = torch.zeros((3,) + shape, dtype=torch.float64, device="cuda")
    C = torch.ones((2,) + shape, dtype=torch.float64, device="cuda")
    with pytest.raises(el.AlamoError, match="exceeds the scratch"):
        op.newton_rhs(C, z, z, out=torch.empty_like(z), noalloc=True)
