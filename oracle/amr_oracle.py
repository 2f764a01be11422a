"""CPU oracle of the composite (AMR) solve: nested single-box patches,
refinement ratio 2, nodal, "ghostnodes" coarse-fine treatment.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ to check alamo_amr_solve.

Restates, in numpy over the C++ oracle's Fapply/Jacobi/residual/MG cycle:
  - the patch boundary table: the parent's on domain faces, displacement on
    every region touching a face inside the domain (ghost layer);
  - trilinear interpolation onto the patch (Operator.cpp:365-434 weights);
  - FillPatch of a regridded patch (src/Integrator/Integrator.cpp:394-478):
    the old patch's unknowns where they reach, interpolation elsewhere;
  - averageDownSolutionRHS, nodal injection (src/Operator/Operator.cpp:436-448);
  - reflux (Operator.cpp:526-633): the parent residual at the covered and
    coarse-fine nodes is replaced by the full-weighting restriction of the
    patch residual (domain-face weights at domain faces);
  - the composite inf-norm with strictly covered parent nodes masked
    (fixUpResidualMask, Operator.cpp:230-251, keeps coarse-fine nodes);
  - one iteration: pre-smooth patches finest first, reflux, one MG cycle on
    the base level, interpolate corrections up, post-smooth, add, average
    down, refresh ghosts.
Parity of the coarse-fine treatment with AMReX's MLMG is unpinned (AMReX is
absent); this restatement pins the device implementation of the same steps.
"""
from __future__ import annotations

import numpy as np

from alamo_amd import bc as B

from . import oracle as O


class Level:
    """One AMR level: C (ncomp, nz, ny, nx), its boundary table, cell size,
    global node index of node 0 (x, y, z), covered parent cells, ghost faces,
    domain cells at this level."""

    def __init__(self, C, sym, bc, dx, origin=(0, 0, 0), clo=None, chi=None, gl=(0, 0, 0), gh=(0, 0, 0),
                 dom=None):
        self.C, self.sym, self.bc, self.dx = C, sym, bc, dx
        self.origin, self.clo, self.chi, self.gl, self.gh = origin, clo, chi, gl, gh
        self.shape = C.shape[1:]
        self.dom = dom
        self.diag = O.diagonal(C, sym, bc, dx, self.shape)


def patch_geometry(parent_dom, clo, chi):
    """(origin, gl, gh, n_cell) of the patch over parent cells [clo, chi]."""
    gl = tuple(int(clo[a] > 0) for a in range(3))
    gh = tuple(int(chi[a] < parent_dom[a]) for a in range(3))
    origin = tuple(2 * clo[a] - gl[a] for a in range(3))
    n_cell = tuple(2 * (chi[a] - clo[a]) + gl[a] + gh[a] for a in range(3))
    return origin, gl, gh, n_cell


def patch_bc(parent_bc, gl, gh):
    inner = {"xlo": gl[0], "ylo": gl[1], "zlo": gl[2], "xhi": gh[0], "yhi": gh[1], "zhi": gh[2]}
    out = B.Constant()
    for r, name in enumerate(B.FACES):
        parts = [name[i:i + 3] for i in range(0, len(name), 3)]
        if any(inner[f] for f in parts):
            out.type[r, :] = B.DISPLACEMENT
            out.val[r, :] = 0.0
        else:
            out.type[r] = parent_bc.type[r]
            out.val[r] = parent_bc.val[r]
    return out


def ghost_mask(shape, gl, gh):
    nz, ny, nx = shape
    m = np.zeros(shape, dtype=bool)
    if gl[0]: m[:, :, 0] = True
    if gh[0]: m[:, :, nx - 1] = True
    if gl[1]: m[:, 0, :] = True
    if gh[1]: m[:, ny - 1, :] = True
    if gl[2]: m[0, :, :] = True
    if gh[2]: m[nz - 1, :, :] = True
    return m


def interp(c, parent, lev):
    """Trilinear interpolation of the parent field onto every node of lev."""
    nz, ny, nx = lev.shape
    out = np.zeros((3,) + lev.shape)
    ox, oy, oz = lev.origin
    px, py, pz = parent.origin
    gx, gy, gz = ox + np.arange(nx), oy + np.arange(ny), oz + np.arange(nz)

    def axis(g, p):
        lo, hi = (g >> 1) - p, ((g + 1) >> 1) - p
        return lo, hi

    xl, xh = axis(gx, px)
    yl, yh = axis(gy, py)
    zl, zh = axis(gz, pz)
    for zi in (zl, zh):
        for yi in (yl, yh):
            for xi in (xl, xh):
                out += c[:, zi[:, None, None], yi[None, :, None], xi[None, None, :]]
    return out / 8.0


def fill_patch(c, parent, lev, old=None, old_u=None):
    """Integrator::FillPatch for a nodal field after a regrid
    (src/Integrator/Integrator.cpp:394-478; FillPatchTwoLevels with
    node_bilinear_interp): the new patch lev takes the old patch's values at
    the nodes its unknowns (global nodes 2 clo .. 2 chi) hold, the parent
    field interpolated trilinearly elsewhere (everywhere without an old
    patch).  parent, lev, old: objects with origin, shape (and clo, chi)."""
    out = interp(c, parent, lev)
    if old is None:
        return out
    nz, ny, nx = lev.shape
    g = [lev.origin[a] + np.arange(n) for a, n in enumerate((nx, ny, nz))]
    inside = [(2 * old.clo[a] <= g[a]) & (g[a] <= 2 * old.chi[a]) for a in range(3)]
    if not all(m.any() for m in inside):
        return out
    sel = [np.nonzero(m)[0] for m in inside]
    src = [g[a][sel[a]] - old.origin[a] for a in range(3)]
    out[:, sel[2][:, None, None], sel[1][None, :, None], sel[0][None, None, :]] = \
        old_u[:, src[2][:, None, None], src[1][None, :, None], src[0][None, None, :]]
    return out


def inject(f, lev, c, parent):
    ox, oy, oz = lev.origin
    px, py, pz = parent.origin
    (x0, y0, z0), (x1, y1, z1) = lev.clo, lev.chi
    I, J, K = np.arange(x0, x1 + 1), np.arange(y0, y1 + 1), np.arange(z0, z1 + 1)
    c[:, (K - pz)[:, None, None], (J - py)[None, :, None], (I - px)[None, None, :]] = \
        f[:, (2 * K - oz)[:, None, None], (2 * J - oy)[None, :, None], (2 * I - ox)[None, None, :]]


def reflux(rf, lev, rc, parent):
    """Operator::reflux (src/Operator/Operator.cpp:526-633): the parent
    residual at the covered and coarse-fine nodes [clo, chi] becomes the nodal
    full weighting of the patch residual (ghost rows zeroed by the caller),
    with the domain-face rules of the MG restriction on domain faces.  Computed
    as the MG restriction of the patch residual extended by zero over a box
    one parent node wider than [clo, chi] on every side inside the domain;
    reflux_loops states the same per node."""
    lo = [2 * lev.clo[a] - 2 if lev.clo[a] > 0 else 0 for a in range(3)]
    hi = [2 * lev.chi[a] + 2 if lev.chi[a] < parent.dom[a] else 2 * parent.dom[a] for a in range(3)]
    buf = np.zeros((rf.shape[0], hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1))
    o = [lev.origin[a] - lo[a] for a in range(3)]
    nz, ny, nx = rf.shape[1:]
    buf[:, o[2]:o[2] + nz, o[1]:o[1] + ny, o[0]:o[0] + nx] = rf
    c = O.restrict(buf)
    cl = [lev.clo[a] - lo[a] // 2 for a in range(3)]
    n = [lev.chi[a] - lev.clo[a] + 1 for a in range(3)]
    p = [lev.clo[a] - parent.origin[a] for a in range(3)]
    rc[:, p[2]:p[2] + n[2], p[1]:p[1] + n[1], p[0]:p[0] + n[0]] = \
        c[:, cl[2]:cl[2] + n[2], cl[1]:cl[1] + n[1], cl[0]:cl[0] + n[0]]


def reflux_loops(rf, lev, rc, parent):
    """reflux, node by node: the 27-point weights inside the domain, the
    1D weights (1/2, 1, 1/2) normalised over the directions a domain face does
    not pin."""
    ox, oy, oz = lev.origin
    px, py, pz = parent.origin
    dnx, dny, dnz = parent.dom
    (x0, y0, z0), (x1, y1, z1) = lev.clo, lev.chi
    w1 = {-1: 0.5, 0: 1.0, 1: 0.5}
    for K in range(z0, z1 + 1):
        for J in range(y0, y1 + 1):
            for I in range(x0, x1 + 1):
                i, j, k = 2 * I - ox, 2 * J - oy, 2 * K - oz
                bi, bj, bk = I in (0, dnx), J in (0, dny), K in (0, dnz)
                # full weighting over the directions not pinned by a domain face
                dxs = [0] if bi else [-1, 0, 1]
                dys = [0] if bj else [-1, 0, 1]
                dzs = [0] if bk else [-1, 0, 1]
                norm = sum(w1[a] for a in dxs) * sum(w1[a] for a in dys) * sum(w1[a] for a in dzs)
                v = 0.0
                for dz in dzs:
                    for dy in dys:
                        for dx in dxs:
                            v = v + w1[dx] * w1[dy] * w1[dz] * rf[:, k + dz, j + dy, i + dx]
                rc[:, K - pz, J - py, I - px] = v / norm


def covered_mask(parent, child):
    nz, ny, nx = parent.shape
    px, py, pz = parent.origin
    I, J, K = px + np.arange(nx), py + np.arange(ny), pz + np.arange(nz)
    (x0, y0, z0), (x1, y1, z1) = child.clo, child.chi
    return (((K > z0) & (K < z1))[:, None, None] & ((J > y0) & (J < y1))[None, :, None]
            & ((I > x0) & (I < x1))[None, None, :])


def residuals(levels, xs, bs):
    n = len(levels)
    rs = [None] * n
    for l in range(n - 1, -1, -1):
        lv = levels[l]
        r = O.residual(xs[l], bs[l], lv.C, lv.sym, lv.bc, lv.dx)
        if l > 0:
            r[:, ghost_mask(lv.shape, lv.gl, lv.gh)] = 0.0
        if l + 1 < n:
            reflux(rs[l + 1], levels[l + 1], r, lv)
        rs[l] = r
    return rs


def masked_norm(levels, vs):
    m = 0.0
    for l, lv in enumerate(levels):
        v = vs[l].copy()
        if l + 1 < len(levels):
            v[:, covered_mask(lv, levels[l + 1])] = 0.0
        if l > 0:
            v[:, ghost_mask(lv.shape, lv.gl, lv.gh)] = 0.0
        m = max(m, float(np.abs(v).max()))
    return m


def sync(levels, xs):
    for l in range(len(levels) - 1, 0, -1):
        inject(xs[l], levels[l], xs[l - 1], levels[l - 1])
    for l in range(1, len(levels)):
        g = ghost_mask(levels[l].shape, levels[l].gl, levels[l].gh)
        xs[l][:, g] = interp(xs[l - 1], levels[l - 1], levels[l])[:, g]


def jacobi(lv, e, r, sweeps, omega=2.0 / 3.0):
    """sweeps damped Jacobi sweeps (sweeps even: sweeps/2 Fsmooth calls)."""
    return O.fsmooth(e, r, lv.C, lv.sym, lv.bc, lv.dx, lv.diag, omega=omega, nsmooth=sweeps // 2)


def amr_solve(levels, xs, bs, tol_rel, tol_abs=0.0, max_iter=100, nu1=2, nu2=2, cycle="F"):
    xs = [x.copy() for x in xs]
    n = len(levels)
    sync(levels, xs)
    rs = residuals(levels, xs, bs)
    r0, bn = masked_norm(levels, rs), masked_norm(levels, bs)
    target = max(tol_abs, max(tol_rel, 1e-16) * max(r0, bn))
    nrm, it = r0, 0
    hist = [r0]
    while nrm > target and it < max_iter:
        es = [None] * n
        for l in range(n - 1, 0, -1):
            lv = levels[l]
            g = ghost_mask(lv.shape, lv.gl, lv.gh)
            e = np.zeros((3,) + lv.shape)
            r = rs[l]
            r[:, g] = e[:, g]
            e = jacobi(lv, e, r, 2 * nu1)
            rr = O.residual(e, r, lv.C, lv.sym, lv.bc, lv.dx)
            rr[:, g] = 0.0
            reflux(rr, lv, rs[l - 1], levels[l - 1])
            es[l] = e
        b0 = levels[0]
        es[0], _ = O.mlmg_solve(rs[0], b0.C, b0.sym, b0.bc, b0.dx, x0=np.zeros_like(rs[0]), tol_rel=1e-300,
                                tol_abs=0.0, max_iter=1, nu1=nu1, nu2=nu2, cycle=cycle)
        for l in range(1, n):
            lv = levels[l]
            g = ghost_mask(lv.shape, lv.gl, lv.gh)
            es[l] = es[l] + interp(es[l - 1], levels[l - 1], lv)
            rs[l][:, g] = es[l][:, g]
            es[l] = jacobi(lv, es[l], rs[l], 2 * nu2)
        for l in range(n):
            xs[l] = xs[l] + es[l]
        sync(levels, xs)
        rs = residuals(levels, xs, bs)
        nrm = masked_norm(levels, rs)
        it += 1
        hist.append(nrm)
    return xs, dict(iters=it, res=nrm, res0=r0, rhsnorm=bn, hist=hist, status=0 if nrm <= target else 1)
