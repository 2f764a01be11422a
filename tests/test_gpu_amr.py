"""GPU tests of the composite AMR solve (alamo_amr_solve, alamo_amr_patch_create)
against the CPU restatement oracle/amr_oracle.py, and of what refinement buys:
the composite solution approaches the uniformly refined one.

Parity bar: the device and the oracle run the same iterations on the same
inputs; after a fixed number of iterations each level's solution agrees to
1e-9 relative L2 (FP64, rounding order only)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from alamo_amd import elastic as el  # noqa: E402
from alamo_amd import model as M  # noqa: E402
from alamo_amd import problems as P  # noqa: E402
from oracle import amr_oracle as AO  # noqa: E402
from oracle import oracle as O  # noqa: E402

SPHERE = [((0.5, 0.5, 0.5), 0.2, 10.0)]


def hierarchy(n0, boxes, spheres=SPHERE, bc=None, width=0.0):
    """Device operators, right-hand sides and oracle levels of a base grid
    with nested patches boxes = [(clo, chi), ...] (parent-level cells)."""
    import alamo_amd.bc as B
    bc = bc or B.Constant()
    ops, bs, lv_or = [], [], []
    base = el.ElasticOperator(n0, 1.0 / n0, M.ISOTROPIC)
    base.SetBC(bc)
    ops.append(base)
    for clo, chi in boxes:
        ops.append(ops[-1].refine(clo, chi))
    dom = (n0, n0, n0)
    for l, op in enumerate(ops):
        shape = op.level_shape(0)
        C, F0 = P.inclusions_on_box(op.origin, shape, op.dx[0], spheres, width=width)
        Cd = torch.from_numpy(C).cuda()
        op.SetModel(Cd)
        b = op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda"))
        rhs = op.newton_rhs(Cd, torch.zeros_like(b), b, F0=torch.from_numpy(F0).cuda())
        bs.append(rhs)
        if l == 0:
            lv_or.append(AO.Level(C, M.ISOTROPIC, bc, op.dx[0], dom=dom))
        else:
            clo, chi = boxes[l - 1]
            origin, gl, gh, n_cell = AO.patch_geometry(dom, clo, chi)
            dom = tuple(2 * d for d in dom)
            assert tuple(origin) == tuple(op.origin) and tuple(gl) == tuple(op.ghost_lo)
            assert tuple(n_cell) == tuple(s - 1 for s in shape[::-1])
            lv_or.append(AO.Level(C, M.ISOTROPIC, AO.patch_bc(bc, gl, gh), op.dx[0], origin=origin, clo=clo,
                                  chi=chi, gl=gl, gh=gh, dom=dom))
    return ops, bs, lv_or


@pytest.mark.parametrize("boxes", [
    [((4, 4, 4), (12, 12, 12))],
    [((0, 2, 4), (10, 14, 12))],                                   # touches the domain at x lo
    [((2, 2, 2), (14, 14, 14)), ((8, 8, 8), (20, 20, 20))],       # three AMR levels
])
def test_composite_iterations_match_oracle(boxes):
    ops, bs, lv = hierarchy(16, boxes)
    xs = [torch.zeros_like(b) for b in bs]
    iters = 4
    info = el.amr_solve(ops, xs, bs, 1e-300, 0.0, max_iter=iters, abort_on_fail=False)
    assert info.iters == iters
    bh = [b.cpu().numpy() for b in bs]
    xo, io = AO.amr_solve(lv, [np.zeros_like(b) for b in bh], bh, 1e-300, max_iter=iters)
    for l in range(len(ops)):
        g = xs[l].cpu().numpy()
        e = np.linalg.norm(g - xo[l]) / np.linalg.norm(xo[l])
        print(f"level {l}: rel L2 device vs oracle {e:.2e}")
        assert e < 1e-9
    assert np.isclose(info.res, io["res"], rtol=1e-6)


def test_composite_solve_converges_and_refinement_helps():
    """Base 17^3 with a patch around the inclusion: the composite solution at
    the patch nodes is closer to the uniform 33^3 solution than the base-grid
    solution interpolated there."""
    boxes = [((3, 3, 3), (13, 13, 13))]
    ops, bs, lv = hierarchy(16, boxes)
    xs = [torch.zeros_like(b) for b in bs]
    # V-cycle: with the F-cycle the reference's Jacobi diverges past ~1e-9 at
    # the E=10 interface, on one level and on the composite alike
    info = el.amr_solve(ops, xs, bs, 1e-10, 0.0, max_iter=200, cycle="V")
    print(f"composite: {info.iters} iterations, |r| {info.res:.3e}")
    assert info.status == 0
    # uniform fine reference and base-only solve
    fine, fb, _ = hierarchy(32, [])
    xf = torch.zeros_like(fb[0])
    el.Linear(fine[0], cycle="V").solve(xf, fb[0], 1e-10, 0.0)
    base, bb, _ = hierarchy(16, [])
    xb = torch.zeros_like(bb[0])
    el.Linear(base[0], cycle="V").solve(xb, bb[0], 1e-10, 0.0)
    p = ops[1]
    o = p.origin
    nz, ny, nx = p.level_shape(0)
    ref = xf[:, o[2]:o[2] + nz, o[1]:o[1] + ny, o[0]:o[0] + nx]
    # base solution interpolated onto the patch nodes
    lvb = AO.Level(bb[0].cpu().numpy()[:2] * 0 + 1, M.ISOTROPIC, fine[0]._bc, 1 / 16, dom=(16, 16, 16))
    lvb.origin = (0, 0, 0)
    lvp = AO.Level(np.ones((2, nz, ny, nx)), M.ISOTROPIC, fine[0]._bc, 1 / 32, origin=o)
    xbi = AO.interp(xb.cpu().numpy(), lvb, lvp)
    e_amr = float(torch.linalg.norm(xs[1] - ref) / torch.linalg.norm(ref))
    e_base = float(np.linalg.norm(xbi - ref.cpu().numpy()) / np.linalg.norm(ref.cpu().numpy()))
    print(f"patch nodes: AMR vs uniform fine {e_amr:.3e}, base grid vs uniform fine {e_base:.3e}")
    assert e_amr < 0.5 * e_base


def test_bad_patches_are_rejected():
    base = el.ElasticOperator(16, 1 / 16, M.ISOTROPIC)
    with pytest.raises(el.AlamoError):
        base.refine((4, 4, 4), (17, 12, 12))      # beyond the domain
    p = base.refine((4, 4, 4), (12, 12, 12))
    with pytest.raises(el.AlamoError):
        p.refine((8, 8, 8), (20, 20, 20))          # ghost layer outside the parent's unknowns


def _eta_of(spheres, width):
    return lambda op: torch.from_numpy(P.sphere_indicators(op.origin, op.level_shape(0), op.dx[0], spheres,
                                                           width)).cuda()


def _device_boxes(n0, spheres, width, levels=3):
    base = el.ElasticOperator(n0, 1.0 / n0, M.ISOTROPIC)
    return [(op.clo, op.chi) for op in el.amr_levels(base, _eta_of(spheres, width), levels=levels)[1:]]


def _oracle_boxes(n0, spheres, width, levels=3, threshold=0.1, buffer=2):
    """The same hierarchy from the oracle's tags: each level's indicator at
    its handle's nodes, the tags' box clipped to the level's unknowns less one
    node where they meet an inner face (the nesting a patch needs)."""
    boxes, origin, gl, gh, dom = [], (0, 0, 0), (0, 0, 0), (0, 0, 0), (n0,) * 3
    n_cell = (n0,) * 3
    while len(boxes) < levels - 1:
        dx = 1.0 / dom[0]
        shape = tuple(n_cell[a] + 1 for a in (2, 1, 0))
        eta = P.sphere_indicators(origin, shape, dx, spheres, width)
        box, n = O.tag_box(eta, (dx,) * 3, threshold, buffer)
        if box is None:
            break
        clo, chi = [], []
        for a in range(3):
            ulo, uhi = origin[a] + gl[a], origin[a] + n_cell[a] - gh[a]
            clo.append(max(box[0][a] + origin[a], ulo + (ulo > 0)))
            chi.append(min(box[1][a] + origin[a], uhi - (uhi < dom[a])))
        if any(chi[a] - clo[a] < 2 for a in range(3)):
            break
        boxes.append((tuple(clo), tuple(chi)))
        origin, gl, gh, n_cell = AO.patch_geometry(dom, clo, chi)
        dom = tuple(2 * d for d in dom)
    return boxes


def test_three_level_random_inclusions_fp64_and_fp32():
    """Configuration of the 3-level AMR case at a 33^3 base: three seeded
    inclusions (r ~ U[0.05, 0.15], E ~ U[2, 20]) with interfaces two base
    cells wide, patches from the indicator-gradient tags; the composite
    V-cycle solve converges in FP64 and with FP32 base cycles, and the two
    agree.  (With sharp 20x jumps the strong-form operator loses ellipticity
    and the reference's multigrid diverges on one level already.)"""
    import time
    spheres = P.random_spheres(3, seed=0)
    w = 2.0 / 32
    boxes = _device_boxes(32, spheres, w)
    assert len(boxes) == 2
    ops, bs, _ = hierarchy(32, boxes, spheres=spheres, width=w)
    res = {}
    for prec in ("fp64", "fp32"):
        xs = [torch.zeros_like(b) for b in bs]
        torch.cuda.synchronize()
        t = time.perf_counter()
        info = el.amr_solve(ops, xs, bs, 1e-8, 0.0, max_iter=300, cycle="V", precision=prec)
        torch.cuda.synchronize()
        res[prec] = (xs, info.iters, time.perf_counter() - t)
        print(f"{prec}: {info.iters} composite V-cycles, {res[prec][2] * 1e3:.1f} ms, |r| {info.res:.3e}")
        assert info.status == 0
    for l in range(3):
        a, b = res["fp32"][0][l], res["fp64"][0][l]
        e = float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
        print(f"level {l}: FP32-refined vs FP64 rel L2 {e:.2e}")
        assert e < 1e-6


@pytest.mark.parametrize("ncomp", [1, 3])
def test_refinement_tagging_matches_oracle(ncomp):
    """alamo_el_tag_bbox (src/Integrator/Mechanics.H:298-313) against the
    oracle: a smooth random field plus sharp and diffuse sphere indicators on
    a 25 x 19 x 31-node box, tagged cells counted and bounded exactly (no node
    sits within 1e-9 of the threshold); none tagged above every gradient."""
    from alamo_amd import elastic as el
    rng = np.random.default_rng(60 + ncomp)
    n_cell = (24, 18, 30)
    shape = (31, 19, 25)
    dx = (1.0 / 24, 0.75 / 18, 1.25 / 30)
    z, y, x = np.meshgrid(np.arange(shape[0]) * dx[2], np.arange(shape[1]) * dx[1], np.arange(shape[2]) * dx[0],
                          indexing="ij")
    fields = [0.02 * np.sin(3 * x + 2 * y) * np.cos(4 * z),
              ((x - 0.4) ** 2 + (y - 0.3) ** 2 + (z - 0.7) ** 2 < 0.04).astype(np.float64),
              0.5 * (1 - np.tanh((np.sqrt((x - 0.6) ** 2 + (y - 0.5) ** 2 + (z - 0.3) ** 2) - 0.15) / 0.05))]
    eta = np.stack([fields[(q + ncomp) % 3] + 1e-3 * rng.standard_normal(shape) for q in range(ncomp)])
    op = el.ElasticOperator(n_cell, dx, M.ISOTROPIC)
    h = np.asarray(dx)
    dxn = np.sqrt((h * h).sum())
    for thr, buf in ((0.1, 2), (0.02, 0), (0.35, 3)):
        for q in range(ncomp):
            g = [np.gradient(eta[q], h[a], axis=2 - a, edge_order=1) for a in range(3)]
            v = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2) * dxn
            assert np.abs(v - thr).min() > 1e-9 * thr
        assert op.tag_box(torch.from_numpy(eta).cuda(), thr, buf) == O.tag_box(eta, dx, thr, buf)
    assert op.tag_box(torch.from_numpy(eta).cuda(), 1e6, 2) == (None, 0)


def test_device_tagged_hierarchy_matches_oracle_tags():
    """amr_levels (device tags, alamo_el_tag_bbox on every level's indicator
    field, then the clipped box) builds the same three levels as the oracle's
    tags (src/Integrator/Mechanics.H:298-313) on the 33^3-base inclusion case
    and on a sharp single sphere."""
    for n0, spheres, width in ((32, P.random_spheres(3, seed=0), 2.0 / 32), (16, SPHERE, 0.0),
                               (32, P.random_spheres(3, seed=7), 2.0 / 32)):
        dev, ora = _device_boxes(n0, spheres, width), _oracle_boxes(n0, spheres, width)
        print(n0, dev)
        assert dev == ora and len(dev) >= 1


@pytest.mark.parametrize("with_old", [False, True])
def test_fillpatch_matches_oracle(with_old):
    """alamo_amr_fillpatch (Integrator::FillPatch / FillCoarsePatch,
    src/Integrator/Integrator.cpp:394-478) against the oracle on random fields:
    a level-1 patch filled from the base and from an overlapping old level-1
    patch, then a level-2 patch of the new level 1 filled from an old level-2
    patch whose parent was the old level 1.  Copied nodes agree bit for bit,
    interpolated ones to 1e-15 (the oracle sums repeated corners)."""
    from types import SimpleNamespace as NS
    rng = np.random.default_rng(8)
    base = el.ElasticOperator((16, 12, 20), 1.0 / 16, M.ISOTROPIC)
    new1, old1 = base.refine((2, 1, 3), (11, 9, 14)), base.refine((5, 0, 8), (15, 7, 19))
    new2, old2 = new1.refine((6, 4, 8), (18, 14, 26)), old1.refine((12, 3, 18), (26, 12, 36))

    def ns(op):
        return NS(origin=op.origin, shape=op.level_shape(0), clo=op.clo, chi=op.chi)

    def field(op):
        return rng.standard_normal((3,) + op.level_shape(0))

    c0, o1, o2 = field(base), field(old1), field(old2)
    par0 = NS(origin=(0, 0, 0), shape=base.level_shape(0))
    f1 = new1.fill_patch(torch.from_numpy(c0).cuda(), old1 if with_old else None,
                         torch.from_numpy(o1).cuda() if with_old else None)
    r1 = AO.fill_patch(c0, par0, ns(new1), ns(old1) if with_old else None, o1 if with_old else None)
    f2 = new2.fill_patch(f1, old2 if with_old else None, torch.from_numpy(o2).cuda() if with_old else None)
    r2 = AO.fill_patch(f1.cpu().numpy(), ns(new1), ns(new2), ns(old2) if with_old else None,
                       o2 if with_old else None)
    for f, r in ((f1, r1), (f2, r2)):
        d = np.abs(f.cpu().numpy() - r)
        print(f"max |device - oracle| {d.max():.2e}, exact at {np.mean(d == 0):.3f} of the nodes")
        assert d.max() <= 1e-15 * np.abs(r).max()
    if with_old:  # nodes held by the old patches are copied exactly
        for f, new, old, o in ((f1, new1, old1, o1), (f2, new2, old2, o2)):
            lo = [max(new.origin[a], 2 * old.clo[a]) for a in range(3)]
            hi = [min(new.origin[a] + new.level_shape(0)[2 - a] - 1, 2 * old.chi[a]) for a in range(3)]
            sn = tuple(slice(lo[a] - new.origin[a], hi[a] - new.origin[a] + 1) for a in (2, 1, 0))
            so = tuple(slice(lo[a] - old.origin[a], hi[a] - old.origin[a] + 1) for a in (2, 1, 0))
            assert np.array_equal(f.cpu().numpy()[(slice(None),) + sn], o[(slice(None),) + so])
    with pytest.raises(el.AlamoError):
        new2.fill_patch(f1, old1, torch.from_numpy(o1).cuda())   # old patch of another level


def test_regrid_after_the_inclusions_move():
    """The adaptive flow on the 33^3-base three-inclusion case: device tags
    build three levels, the composite solve converges; the inclusions move by
    one base cell, amr_regrid retags and fills the new levels from the old
    solution, and the composite solve from the filled fields converges to the
    solution reached from zero (1e-6 relative on every level)."""
    n0, w = 32, 2.0 / 32
    spheres = P.random_spheres(3, seed=0)
    ops = el.amr_levels(el.ElasticOperator(n0, 1.0 / n0, M.ISOTROPIC), _eta_of(spheres, w), levels=3)
    bs = _models(ops, spheres, w)
    xs = [torch.zeros_like(b) for b in bs]
    info = el.amr_solve(ops, xs, bs, 1e-8, 0.0, max_iter=300, cycle="V")
    assert info.status == 0
    moved = [((c[0] + 1.0 / n0, c[1], c[2]), r, E) for c, r, E in spheres]
    ops2, xs2 = el.amr_regrid(ops, xs, _eta_of(moved, w), levels=3)
    assert len(ops2) == 3 and [(o.clo, o.chi) for o in ops2[1:]] == _oracle_boxes(n0, moved, w)
    assert [(o.clo, o.chi) for o in ops2[1:]] != [(o.clo, o.chi) for o in ops[1:]]
    bs2 = _models(ops2, moved, w)
    sol = {}
    for name, x0 in (("filled", xs2), ("zero", [torch.zeros_like(b) for b in bs2])):
        x = [v.clone() for v in x0]
        info = el.amr_solve(ops2, x, bs2, 1e-8, 0.0, max_iter=300, cycle="V")
        print(f"from {name}: {info.iters} composite V-cycles, |r0| {info.res0:.2e}, |b| {info.rhsnorm:.2e}, "
              f"|r| {info.res:.2e}")
        assert info.status == 0
        sol[name] = x
    for a, b in zip(sol["filled"], sol["zero"]):
        assert float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) < 1e-6


def _models(ops, spheres, width):
    bs = []
    for op in ops:
        shape = op.level_shape(0)
        C, F0 = P.inclusions_on_box(op.origin, shape, op.dx[0], spheres, width=width)
        Cd = torch.from_numpy(C).cuda()
        op.SetModel(Cd)
        b0 = op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda"))
        bs.append(op.newton_rhs(Cd, torch.zeros_like(b0), b0, F0=torch.from_numpy(F0).cuda()))
    return bs


def test_plotfile_of_a_composite_solution(tmp_path):
    """The solved three-level hierarchy written as the reference's node and
    cell plotfiles (eta, disp, stress per level, stress from alamo_el_stress
    on each handle) reads back with the solution on every level's unknowns."""
    from alamo_amd import plotfile as PF
    n0, w = 16, 2.0 / 16
    spheres = [((0.5, 0.5, 0.5), 0.2, 10.0)]
    ops = el.amr_levels(el.ElasticOperator(n0, 1.0 / n0, M.ISOTROPIC), _eta_of(spheres, w), levels=3)
    bs = _models(ops, spheres, w)
    xs = [torch.zeros_like(b) for b in bs]
    assert el.amr_solve(ops, xs, bs, 1e-8, 0.0, max_iter=300, cycle="V").status == 0
    eta, sig = [], []
    for op, x in zip(ops, xs):
        C, F0 = P.inclusions_on_box(op.origin, op.level_shape(0), op.dx[0], spheres, width=w)
        sig.append(op.stress(torch.from_numpy(C).cuda(), x, F0=torch.from_numpy(F0).cuda()))
        eta.append(_eta_of(spheres, w)(op))
    paths = PF.write_mechanics(str(tmp_path / "out"), 0, ops, [("eta", eta), ("disp", xs), ("stress", sig)])
    names, _, levels = PF.read(paths[0])
    assert names[:4] == ["eta", "disp_x", "disp_y", "disp_z"] and len(levels) == len(ops) == 3
    for (lo, hi, data), op, x in zip(levels, ops, xs):
        gl, gh = op.ghost_lo, op.ghost_hi
        v = x.cpu().numpy()
        nz, ny, nx = v.shape[1:]
        assert lo == tuple(op.origin[a] + gl[a] for a in range(3))
        assert np.array_equal(data[1:4], v[:, gl[2]:nz - gh[2], gl[1]:ny - gh[1], gl[0]:nx - gh[0]])


def test_uniaxial_tension_deck_three_levels_closed_form():
    """tests/UniaxialTension in 3D (input:37-99: base 32^3 cells over
    [-8, 8]^3, explicit mesh of the middle half on each of two levels, E = 210,
    nu = 0.3): the device's composite FP64 solve reproduces the checker's
    closed form u_x = (x + 8) 0.1/16, sigma_xx = E 0.1/16 (test:27-28,
    tolerance 1e-8 on the squared relative error) on every level, and so do
    the FP32 cycles."""
    from test_oracle_closed_forms import uniaxial_tension_deck_bc, uniaxial_tension_levels
    bc = uniaxial_tension_deck_bc()
    lam, mu = M.lame_from_E_nu(210.0, 0.3)
    eps = 0.1 / 16
    for precision in ("fp64", "fp32"):
        base = el.ElasticOperator(32, 0.5, M.ISOTROPIC)
        base.SetBC(bc)
        ops = [base]
        for clo, chi in uniaxial_tension_levels(32):
            ops.append(ops[-1].refine(clo, chi))
        bs, Cs = [], []
        for op in ops:
            shape = op.level_shape(0)
            C = torch.empty((2,) + shape, dtype=torch.float64, device="cuda")
            C[0], C[1] = lam, mu
            op.SetModel(C)
            Cs.append(C)
            bs.append(op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda")))
        xs = [torch.zeros_like(b) for b in bs]
        info = el.amr_solve(ops, xs, bs, 1e-10, 0.0, max_iter=200, precision=precision)
        assert info.status == 0
        for l, op in enumerate(ops):
            nz, ny, nx = op.level_shape(0)
            h = op.dx[0]
            x = -8.0 + (op.origin[0] + torch.arange(nx, device="cuda", dtype=torch.float64)) * h
            ux = ((x + 8.0) * eps).expand(nz, ny, nx)
            err = float(((xs[l][0] - ux) ** 2).sum() / (ux ** 2).sum())
            sig = op.stress(Cs[l], xs[l])
            serr = float(((sig[0] - 210.0 * eps) ** 2).sum() / ((210.0 * eps) ** 2 * sig[0].numel()))
            print(f"{precision} level {l}: {info.iters} iterations, u_x squared rel error {err:.2e}, "
                  f"sigma_xx {serr:.2e}")
            assert err < 1e-8 and serr < 1e-8


def _multibox(n0, spheres, width, grid_eff=0.7, blocking_factor=4, levels=3):
    """Device hierarchy with several patches per level (clustered tags), each
    patch's model and right-hand side, and the matching oracle levels."""
    import alamo_amd.bc as B
    base = el.ElasticOperator(n0, 1.0 / n0, M.ISOTROPIC)
    base.SetBC(B.Constant())
    ops = el.amr_levels(base, _eta_of(spheres, width), levels=levels, threshold=0.1, buffer=2,
                        blocking_factor=blocking_factor, grid_eff=grid_eff)
    bs, lv = [], []
    for q, op in enumerate(ops):
        shape = op.level_shape(0)
        C, F0 = P.inclusions_on_box(op.origin, shape, op.dx[0], spheres, width=width)
        op.SetModel(torch.from_numpy(C).cuda())
        b = op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda"))
        bs.append(op.newton_rhs(torch.from_numpy(C).cuda(), torch.zeros_like(b), b, F0=torch.from_numpy(F0).cuda()))
        if q == 0:
            lv.append(AO.Level(C, M.ISOTROPIC, B.Constant(), op.dx[0], dom=op.dom))
        else:
            par = next(i for i, o in enumerate(ops) if o is op.parent)
            L = AO.Level(C, M.ISOTROPIC, AO.patch_bc(B.Constant(), op.ghost_lo, op.ghost_hi), op.dx[0],
                         origin=op.origin, clo=op.clo, chi=op.chi, gl=op.ghost_lo, gh=op.ghost_hi, dom=op.dom)
            L.parent = par
            lv.append(L)
    return ops, bs, lv


def test_multibox_hierarchy_matches_oracle_clustering():
    """Several patches per level (alamo_el_tag_cells + alamo_amr_cluster):
    the device-built tree equals the oracle's, from its own tags and its own
    Berger-Rigoutsos, on the three-inclusion case at a 33^3 base and on a
    second seed."""
    for seed in (0, 7):
        spheres = P.random_spheres(3, seed=seed)
        w = 2.0 / 32
        base = el.ElasticOperator(32, 1.0 / 32, M.ISOTROPIC)
        ops = el.amr_levels(base, _eta_of(spheres, w), levels=3, threshold=0.1, buffer=2, blocking_factor=4,
                            grid_eff=0.7)
        dev = [(next(i for i, o in enumerate(ops) if o is op.parent), op.clo, op.chi) for op in ops[1:]]
        ora = [(p, clo, chi) for p, clo, chi, *_ in AO.hierarchy(
            32, lambda o, s, h: P.sphere_indicators(o, s, h, spheres, w), levels=3, threshold=0.1, buffer=2,
            blocking_factor=4, grid_eff=0.7)]
        print(f"seed {seed}: {dev}")
        assert dev == ora
        assert len(dev) > 2                                # several patches on a level


def test_multibox_composite_matches_oracle():
    spheres = P.random_spheres(3, seed=0)
    w = 2.0 / 32
    ops, bs, lv = _multibox(32, spheres, w)
    assert len(ops) > 3
    xs = [torch.zeros_like(b) for b in bs]
    info = el.amr_solve(ops, xs, bs, 1e-8, 0.0, max_iter=300, cycle="V")
    bh = [b.cpu().numpy() for b in bs]
    xo, io = AO.amr_solve(lv, [np.zeros_like(b) for b in bh], bh, 1e-8, max_iter=300, cycle="V")
    print(f"{len(ops)} handles: device {info.iters} composite V-cycles, oracle {io['iters']}")
    assert info.status == 0 and io["status"] == 0 and info.iters == io["iters"]
    for q in range(len(ops)):
        e = np.linalg.norm(xs[q].cpu().numpy() - xo[q]) / np.linalg.norm(xo[q])
        print(f"  handle {q} {ops[q].level_shape(0)}: rel L2 device vs oracle {e:.2e}")
        assert e < 1e-9


def test_patches_of_one_level_must_not_share_nodes():
    base = el.ElasticOperator(16, 1 / 16, M.ISOTROPIC)
    a = base.refine((2, 2, 2), (8, 8, 8))
    b = base.refine((8, 3, 3), (12, 9, 9))      # shares the parent nodes of x = 8 with a
    shape = [op.level_shape(0) for op in (base, a, b)]
    for op, sh in zip((base, a, b), shape):
        op.SetModel(torch.ones((2,) + sh, dtype=torch.float64, device="cuda"))
    xs = [torch.zeros((3,) + sh, dtype=torch.float64, device="cuda") for sh in shape]
    with pytest.raises(el.AlamoError, match="share nodes"):
        el.amr_solve([base, a, b], xs, [x.clone() for x in xs], 1e-8, 0.0)


def test_reflux_entry_point_matches_oracle():
    """alamo_amr_reflux (Operator::reflux, src/Operator/Operator.cpp:526-633)
    against the oracle on random residuals: a patch inside the domain and one
    on two domain faces; the caller's fine residual is not modified."""
    rng = np.random.default_rng(17)
    base = el.ElasticOperator(16, 1 / 16, M.ISOTROPIC)
    for clo, chi in (((3, 4, 2), (11, 12, 9)), ((0, 5, 6), (7, 12, 16))):
        p = base.refine(clo, chi)
        rf = rng.standard_normal((3,) + p.level_shape(0))
        rc = rng.standard_normal((3,) + base.level_shape(0))
        rf_d, rc_d = torch.from_numpy(rf).cuda(), torch.from_numpy(rc.copy()).cuda()
        p.reflux(rf_d, rc_d)
        torch.cuda.synchronize()
        lvp = AO.Level(np.ones((2,) + p.level_shape(0)), M.ISOTROPIC, p._bc, p.dx[0], origin=p.origin, clo=p.clo,
                       chi=p.chi, gl=p.ghost_lo, gh=p.ghost_hi, dom=p.dom)
        lvb = AO.Level(np.ones((2,) + base.level_shape(0)), M.ISOTROPIC, base._bc, base.dx[0], dom=base.dom)
        rz = rf.copy()
        rz[:, AO.ghost_mask(rz.shape[1:], p.ghost_lo, p.ghost_hi)] = 0.0
        want = rc.copy()
        AO.reflux_loops(rz, lvp, want, lvb)
        assert np.abs(rc_d.cpu().numpy() - want).max() <= 1e-14 * np.abs(want).max()
        assert np.array_equal(rf_d.cpu().numpy(), rf)


@pytest.mark.parametrize("boxes", [
    [((2, 2, 2), (14, 14, 14)), ((8, 8, 8), (20, 20, 20))],       # three AMR levels
    [((0, 2, 4), (10, 16, 12))],                                   # on the domain faces x lo and y hi
])
def test_average_down_coeffs_matches_oracle(boxes):
    """alamo_amr_average_down_coeffs (Elastic::averageDownCoeffsDifferentAmrLevels,
    src/Operator/Elastic.cpp:628-745) against the oracle's restatement: the
    parents' stiffness after the averaging, deepest patch first, to rounding,
    and the composite iterations of the averaged hierarchy equal the oracle's."""
    ops, bs, lv = hierarchy(16, boxes, width=1.0 / 16)
    Cs = [torch.from_numpy(l.C.copy()).cuda() for l in lv]
    el.amr_average_down_coeffs(ops, Cs)
    lv_avg = AO.average_down_coeffs(lv)
    for l in range(len(ops)):
        e = np.abs(Cs[l].cpu().numpy() - lv_avg[l].C).max() / np.abs(lv_avg[l].C).max()
        print(f"level {l}: max |device - oracle| / max |C| = {e:.2e}")
        assert e < 1e-15
    changed = np.abs(lv_avg[0].C - lv[0].C).max()
    assert changed > 1e-3  # the averaging did change the parent
    xs = [torch.zeros_like(b) for b in bs]
    iters = 4
    info = el.amr_solve(ops, xs, bs, 1e-300, 0.0, max_iter=iters, abort_on_fail=False)
    bh = [b.cpu().numpy() for b in bs]
    xo, io = AO.amr_solve(lv_avg, [np.zeros_like(b) for b in bh], bh, 1e-300, max_iter=iters)
    for l in range(len(ops)):
        e = np.linalg.norm(xs[l].cpu().numpy() - xo[l]) / np.linalg.norm(xo[l])
        print(f"level {l}: rel L2 device vs oracle {e:.2e}")
        assert e < 1e-9
