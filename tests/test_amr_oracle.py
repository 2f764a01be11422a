"""Pins of the composite-solve oracle (oracle/amr_oracle.py) by mathematics:
trilinear patch interpolation reproduces linear fields and injection inverts
it at the coarse nodes; reflux over a patch that covers the whole domain is
the MG restriction; the patch boundary table marks exactly the regions on
inner faces; and with a patch covering the whole domain the composite solve
converges to the single-level solution of the fine grid."""
import numpy as np

from alamo_amd import bc as B
from alamo_amd import model as M
from alamo_amd import problems as P
from oracle import amr_oracle as AO
from oracle import oracle as O

SPHERE = [((0.5, 0.5, 0.5), 0.2, 10.0)]


def _level(origin, shape, dx, bc, **kw):
    C, _ = P.inclusions_on_box(origin, shape, dx, SPHERE)
    return AO.Level(C, M.ISOTROPIC, bc, dx, origin=origin, **kw)


def test_patch_geometry_and_table():
    origin, gl, gh, n_cell = AO.patch_geometry((16, 16, 16), (0, 3, 5), (8, 16, 9))
    assert gl == (0, 1, 1) and gh == (1, 0, 1)
    assert origin == (0, 5, 9) and n_cell == (17, 27, 10)
    t = AO.patch_bc(B.uniaxial(2, 1e-3), gl, gh)
    F = B.FACE_INDEX
    assert list(t.type[F["ylo"]]) == [B.DISPLACEMENT] * 3 and t.val[F["ylo"], 2] == 0.0
    assert list(t.type[F["xlo"]]) == list(B.uniaxial(2, 1e-3).type[F["xlo"]])   # domain face kept
    assert list(t.type[F["yhi"]]) == list(B.uniaxial(2, 1e-3).type[F["yhi"]])   # domain face kept
    assert list(t.type[F["xloylo"]]) == [B.DISPLACEMENT] * 3                      # edge on an inner face


def test_interp_reproduces_linears_and_inject_inverts_it():
    par = _level((0, 0, 0), (9, 9, 9), 1 / 8, B.Constant(), dom=(8, 8, 8))
    origin, gl, gh, n_cell = AO.patch_geometry((8, 8, 8), (2, 1, 3), (6, 7, 8))
    shape = (n_cell[2] + 1, n_cell[1] + 1, n_cell[0] + 1)
    pat = _level(origin, shape, 1 / 16, AO.patch_bc(B.Constant(), gl, gh), clo=(2, 1, 3), chi=(6, 7, 8), gl=gl,
                 gh=gh, dom=(16, 16, 16))
    z, y, x = np.meshgrid(*[np.arange(9.0)] * 3, indexing="ij")
    c = np.stack([1 + 2 * x - y, 3 * z + 0.5 * x, x + y + z])
    f = AO.interp(c, par, pat)
    fz, fy, fx = np.meshgrid(*[(origin[a] + np.arange(shape[2 - a])) / 2.0 for a in (2, 1, 0)], indexing="ij")
    assert np.allclose(f, np.stack([1 + 2 * fx - fy, 3 * fz + 0.5 * fx, fx + fy + fz]))
    c2 = np.zeros_like(c)
    AO.inject(f, pat, c2, par)
    I = np.s_[:, 3:9, 1:8, 2:7]
    assert np.allclose(c2[I], c[I]) and np.all(c2[:, :3] == 0)


def test_reflux_over_the_whole_domain_is_the_mg_restriction():
    par = _level((0, 0, 0), (9, 9, 9), 1 / 8, B.Constant(), dom=(8, 8, 8))
    pat = _level((0, 0, 0), (17, 17, 17), 1 / 16, B.Constant(), clo=(0, 0, 0), chi=(8, 8, 8), dom=(16, 16, 16))
    rf = np.random.default_rng(0).standard_normal((3, 17, 17, 17))
    rc = np.zeros((3, 9, 9, 9))
    AO.reflux(rf, pat, rc, par)
    assert np.allclose(rc, O.restrict(rf), rtol=1e-14, atol=1e-15)


def test_whole_domain_patch_gives_the_fine_grid_solution():
    bc = B.Constant()
    base = _level((0, 0, 0), (9, 9, 9), 1 / 8, bc, dom=(8, 8, 8))
    pat = _level((0, 0, 0), (17, 17, 17), 1 / 16, bc, clo=(0, 0, 0), chi=(8, 8, 8), dom=(16, 16, 16))
    levels = [base, pat]
    bs = []
    for lv in levels:
        _, F0 = P.inclusions_on_box(lv.origin, lv.shape, lv.dx, SPHERE)
        bs.append(O.newton_rhs(np.zeros((3,) + lv.shape), O.bc_init(lv.shape, bc), lv.C, M.ISOTROPIC, bc, lv.dx,
                               F0=F0))
    xs, info = AO.amr_solve(levels, [np.zeros_like(b) for b in bs], bs, 1e-11, max_iter=60, cycle="V")
    assert info["status"] == 0
    xf, _ = O.mlmg_solve(bs[1], pat.C, M.ISOTROPIC, bc, pat.dx, tol_rel=1e-11, cycle="V")
    assert np.linalg.norm(xs[1] - xf) / np.linalg.norm(xf) < 1e-8
    # the base level holds the injected fine solution
    assert np.allclose(xs[0], xs[1][:, ::2, ::2, ::2])


def test_fill_patch_copies_the_old_unknowns_and_interpolates_the_rest():
    """Regrid FillPatch (src/Integrator/Integrator.cpp:394-478): without an old
    patch it is the interpolation; an old field that is itself the parent's
    interpolation changes nothing; a marked old field lands exactly on the
    overlap of the old unknowns (global nodes 2 clo .. 2 chi per axis, each
    value at its own global node), and every other node keeps the
    interpolated value."""
    from types import SimpleNamespace as NS
    par = NS(origin=(0, 0, 0), shape=(17, 13, 9))                     # 8 x 12 x 16 cells
    dom = (8, 12, 16)

    def patch(clo, chi):
        origin, gl, gh, n_cell = AO.patch_geometry(dom, clo, chi)
        return NS(origin=origin, shape=(n_cell[2] + 1, n_cell[1] + 1, n_cell[0] + 1), clo=clo, chi=chi)

    new, old = patch((1, 2, 3), (6, 9, 12)), patch((3, 0, 6), (8, 5, 15))
    c = np.random.default_rng(5).standard_normal((3,) + par.shape)
    base = AO.interp(c, par, new)
    assert np.array_equal(AO.fill_patch(c, par, new), base)
    assert np.allclose(AO.fill_patch(c, par, new, old, AO.interp(c, par, old)), base, rtol=0, atol=1e-15)
    # mark each old node by its global index: value = gx + 1000 gy + 1e6 gz (+ 1e9 comp)
    oz, oy, ox = np.meshgrid(*[old.origin[a] + np.arange(old.shape[2 - a]) for a in (2, 1, 0)], indexing="ij")
    mark = np.stack([ox + 1e3 * oy + 1e6 * oz + 1e9 * p for p in range(3)]).astype(np.float64)
    f = AO.fill_patch(c, par, new, old, mark)
    nz, ny, nx = np.meshgrid(*[new.origin[a] + np.arange(new.shape[2 - a]) for a in (2, 1, 0)], indexing="ij")
    hit = ((2 * 3 <= nx) & (nx <= 2 * 8) & (2 * 0 <= ny) & (ny <= 2 * 5) & (2 * 6 <= nz) & (nz <= 2 * 15))
    span = [min(new.origin[a] + new.shape[2 - a] - 1, 2 * old.chi[a]) - max(new.origin[a], 2 * old.clo[a]) + 1
            for a in range(3)]
    assert span == [8, 8, 14] and hit.sum() == 8 * 8 * 14
    for p in range(3):
        assert np.array_equal(f[p][hit], (nx + 1e3 * ny + 1e6 * nz + 1e9 * p)[hit])
        assert np.array_equal(f[p][~hit], base[p][~hit])
