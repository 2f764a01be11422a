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
    patch residual (domain-face weights at domain faces); the same
    restriction of the patch's C is averageDownCoeffsDifferentAmrLevels
    (Elastic.cpp:628-745);
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


def average_down_coeffs(levels):
    """Elastic::averageDownCoeffs with average_down_coeffs on
    (src/Operator/Elastic.cpp:605-614,628-745): for every patch, deepest first,
    the parent's C at the covered and coarse-fine nodes [clo, chi] becomes the
    nodal full weighting of the patch's C, ghost layer included, with the
    corner, edge and face weights on domain faces (the reflux restriction on
    the C planes).  Returns new Levels (C and diagonal updated)."""
    import copy
    out = [copy.copy(lv) for lv in levels]
    for lv in out:
        lv.C = lv.C.copy()
    par, kids, order = _tree(out)
    for i in order:
        if par[i] is None:
            continue
        reflux(out[i].C, out[i], out[par[i]].C, out[par[i]])
    for lv in out:
        lv.diag = O.diagonal(lv.C, lv.sym, lv.bc, lv.dx, lv.shape)
    return out


def covered_mask(parent, child):
    nz, ny, nx = parent.shape
    px, py, pz = parent.origin
    I, J, K = px + np.arange(nx), py + np.arange(ny), pz + np.arange(nz)
    (x0, y0, z0), (x1, y1, z1) = child.clo, child.chi
    return (((K > z0) & (K < z1))[:, None, None] & ((J > y0) & (J < y1))[None, :, None]
            & ((I > x0) & (I < x1))[None, None, :])


def _tree(levels):
    """(parent, children, deepest-first order) of a level list: levels[0] is
    the base; a Level with a `parent` index hangs under that level, otherwise
    under the one before it (a chain of single-box levels)."""
    n = len(levels)
    par = [None] + [getattr(levels[i], "parent", None) for i in range(1, n)]
    par = [None] + [i - 1 if par[i] is None else par[i] for i in range(1, n)]
    depth = [0] * n
    for i in range(1, n):
        depth[i] = depth[par[i]] + 1
    kids = [[c for c in range(n) if par[c] == i] for i in range(n)]
    order = sorted(range(n), key=lambda i: -depth[i])
    return par, kids, order


def residuals(levels, xs, bs):
    par, kids, order = _tree(levels)
    rs = [None] * len(levels)
    for l in order:
        lv = levels[l]
        r = O.residual(xs[l], bs[l], lv.C, lv.sym, lv.bc, lv.dx)
        if l > 0:
            r[:, ghost_mask(lv.shape, lv.gl, lv.gh)] = 0.0
        for c in kids[l]:
            reflux(rs[c], levels[c], r, lv)
        rs[l] = r
    return rs


def masked_norm(levels, vs):
    par, kids, order = _tree(levels)
    m = 0.0
    for l, lv in enumerate(levels):
        v = vs[l].copy()
        for c in kids[l]:
            v[:, covered_mask(lv, levels[c])] = 0.0
        if l > 0:
            v[:, ghost_mask(lv.shape, lv.gl, lv.gh)] = 0.0
        m = max(m, float(np.abs(v).max()))
    return m


def sync(levels, xs):
    par, kids, order = _tree(levels)
    for l in order:
        if l > 0:
            inject(xs[l], levels[l], xs[par[l]], levels[par[l]])
    for l in reversed(order):
        if l > 0:
            g = ghost_mask(levels[l].shape, levels[l].gl, levels[l].gh)
            xs[l][:, g] = interp(xs[par[l]], levels[par[l]], levels[l])[:, g]


def jacobi(lv, e, r, sweeps, omega=2.0 / 3.0):
    """sweeps damped Jacobi sweeps (sweeps even: sweeps/2 Fsmooth calls)."""
    return O.fsmooth(e, r, lv.C, lv.sym, lv.bc, lv.dx, lv.diag, omega=omega, nsmooth=sweeps // 2)


def amr_solve(levels, xs, bs, tol_rel, tol_abs=0.0, max_iter=100, nu1=2, nu2=2, cycle="F"):
    """Composite solve over levels[0] (the base) and its patches: a chain of
    single-box levels, or a tree when a Level carries the index of its
    parent (`parent`); patches of one depth share no nodes."""
    xs = [x.copy() for x in xs]
    n = len(levels)
    par, kids, order = _tree(levels)
    sync(levels, xs)
    rs = residuals(levels, xs, bs)
    r0, bn = masked_norm(levels, rs), masked_norm(levels, bs)
    target = max(tol_abs, max(tol_rel, 1e-16) * max(r0, bn))
    nrm, it = r0, 0
    hist = [r0]
    while nrm > target and it < max_iter:
        es = [None] * n
        for l in order:
            if l == 0:
                continue
            lv = levels[l]
            g = ghost_mask(lv.shape, lv.gl, lv.gh)
            e = np.zeros((3,) + lv.shape)
            r = rs[l]
            r[:, g] = e[:, g]
            e = jacobi(lv, e, r, 2 * nu1)
            rr = O.residual(e, r, lv.C, lv.sym, lv.bc, lv.dx)
            rr[:, g] = 0.0
            reflux(rr, lv, rs[par[l]], levels[par[l]])
            es[l] = e
        b0 = levels[0]
        es[0], _ = O.mlmg_solve(rs[0], b0.C, b0.sym, b0.bc, b0.dx, x0=np.zeros_like(rs[0]), tol_rel=1e-300,
                                tol_abs=0.0, max_iter=1, nu1=nu1, nu2=nu2, cycle=cycle)
        for l in reversed(order):
            if l == 0:
                continue
            lv = levels[l]
            g = ghost_mask(lv.shape, lv.gl, lv.gh)
            es[l] = es[l] + interp(es[par[l]], levels[par[l]], lv)
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


# --- Grids from tags (AmrCore::MakeNewGrids with amrex::ClusterList) ---------
def _signature(t, b, a):
    """Tag counts of the planes of box b (lo, hi inclusive, xyz) normal to axis a."""
    sub = t[b[0][2]:b[1][2] + 1, b[0][1]:b[1][1] + 1, b[0][0]:b[1][0] + 1]
    return sub.sum(axis=tuple(ax for ax in (0, 1, 2) if ax != 2 - a)).astype(np.int64)


def _bounding(t, b):
    sub = t[b[0][2]:b[1][2] + 1, b[0][1]:b[1][1] + 1, b[0][0]:b[1][0] + 1]
    k, j, i = np.nonzero(sub)
    if len(i) == 0:
        return None, 0
    lo = (b[0][0] + int(i.min()), b[0][1] + int(j.min()), b[0][2] + int(k.min()))
    hi = (b[0][0] + int(i.max()), b[0][1] + int(j.max()), b[0][2] + int(k.max()))
    return (lo, hi), len(i)


def _berger_rigoutsos(t, box, eff, out):
    """Berger & Rigoutsos (IEEE Trans. SMC 21(5), 1991) restated: shrink to the
    tags, keep at tagged fraction >= eff, else cut at a hole of the signature
    (longest axis first, the hole nearest the middle), else at the inflection
    with the largest jump of the second difference (ties: nearest the middle,
    then the earlier axis of the longest-first order), else halve the longest
    axis; recurse on both halves."""
    b, n = _bounding(t, box)
    if b is None:
        return
    lo, hi = b
    ext = [hi[a] - lo[a] + 1 for a in range(3)]
    if n >= eff * ext[0] * ext[1] * ext[2]:
        out.append(b)
        return
    order = sorted(range(3), key=lambda a: -ext[a])     # stable: lowest axis first among equals
    cut = None
    for a in order:
        s = _signature(t, b, a)
        holes = [m for m in range(1, len(s) - 1) if s[m] == 0]
        if holes:
            m = min(holes, key=lambda m: (abs(2 * m - len(s)), m))
            cut = (a, lo[a] + m)
            break
    if cut is None:
        best = None
        for a in order:
            s = _signature(t, b, a)
            lap = s[:-2] - 2 * s[1:-1] + s[2:]          # lap[m - 1] is at cell m
            for m in range(1, len(s) - 2):
                l0, l1 = lap[m - 1], lap[m]
                if (l0 < 0 < l1) or (l0 > 0 > l1):
                    key = (-abs(int(l1 - l0)), abs(2 * (m + 1) - len(s)))
                    if best is None or key < best[0]:
                        best = (key, a, m + 1)
        if best is not None:
            cut = (best[1], lo[best[1]] + best[2])
    if cut is None:
        a = order[0]
        if ext[a] < 2:
            out.append(b)
            return
        cut = (a, lo[a] + ext[a] // 2)
    a, c = cut
    h1 = list(hi)
    h1[a] = c - 1
    l2 = list(lo)
    l2[a] = c
    _berger_rigoutsos(t, (lo, tuple(h1)), eff, out)
    _berger_rigoutsos(t, (tuple(l2), hi), eff, out)


def cluster_boxes(tags, origin=(0, 0, 0), buffer=1, blocking_factor=8, grid_eff=1.0):
    """Boxes [(clo, chi)] (global cells, chi = last cell + 1) of the next level
    from cell tags (nz, ny, nx) whose cell (0, 0, 0) is global cell `origin`:
    tags grown by `buffer`, coarsened by blocking_factor / 2 on the global grid,
    Berger-Rigoutsos, refined and clipped to the array, boxes closer than one
    cell merged into their bounding box, sorted by (z, y, x) of their corner."""
    t = np.asarray(tags, dtype=bool)
    nz, ny, nx = t.shape
    if buffer > 0:
        g = t.copy()
        for dk in range(-buffer, buffer + 1):
            for dj in range(-buffer, buffer + 1):
                for di in range(-buffer, buffer + 1):
                    src = t[max(0, -dk):nz - max(0, dk), max(0, -dj):ny - max(0, dj), max(0, -di):nx - max(0, di)]
                    g[max(0, dk):nz - max(0, -dk), max(0, dj):ny - max(0, -dj), max(0, di):nx - max(0, -di)] |= src
        t = g
    q = max(1, blocking_factor // 2)
    n = (nx, ny, nz)
    clo = [origin[a] // q for a in range(3)]
    cn = [(origin[a] + n[a] - 1) // q - clo[a] + 1 for a in range(3)]
    c = np.zeros((cn[2], cn[1], cn[0]), dtype=bool)
    k, j, i = np.nonzero(t)
    c[(origin[2] + k) // q - clo[2], (origin[1] + j) // q - clo[1], (origin[0] + i) // q - clo[0]] = True
    coarse = []
    _berger_rigoutsos(c, ((0, 0, 0), (cn[0] - 1, cn[1] - 1, cn[2] - 1)), grid_eff, coarse)
    boxes = []
    for lo, hi in coarse:
        boxes.append([[max((lo[a] + clo[a]) * q, origin[a]) for a in range(3)],
                      [min((hi[a] + clo[a] + 1) * q - 1, origin[a] + n[a] - 1) for a in range(3)]])
    merged = True
    while merged:
        merged = False
        for x in range(len(boxes)):
            for y in range(x + 1, len(boxes)):
                bx, by = boxes[x], boxes[y]
                if all(bx[1][a] + 1 >= by[0][a] and by[1][a] + 1 >= bx[0][a] for a in range(3)):
                    boxes[x] = [[min(bx[0][a], by[0][a]) for a in range(3)], [max(bx[1][a], by[1][a]) for a in range(3)]]
                    del boxes[y]
                    merged = True
                    break
            if merged:
                break
    boxes.sort(key=lambda b: (b[0][2], b[0][1], b[0][0]))
    return [(tuple(b[0]), tuple(v + 1 for v in b[1])) for b in boxes]


def tag_cells(eta, dx, threshold):
    """The cell tags of oracle.tag_box (nz-1, ny-1, nx-1), before the bounding box."""
    eta = np.asarray(eta, dtype=np.float64)
    h = np.broadcast_to(np.asarray(dx, dtype=np.float64), (3,))
    dxn = np.sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])
    tags = np.zeros(eta.shape[1:], dtype=bool)
    for n in range(eta.shape[0]):
        g = [np.gradient(eta[n], h[a], axis=2 - a, edge_order=1) for a in range(3)]
        tags |= np.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) * dxn > threshold
    return tags[:-1, :-1, :-1]


def child_boxes(eta, origin, shape, gl, gh, dom, dx, threshold, buffer, blocking_factor, grid_eff, min_cells=2):
    """The next level's boxes of one handle (ElasticOperator.child_boxes restated):
    its cell tags, only those among the patch's own cells, clustered, each box
    clipped so that the patch and its ghost layer lie on the handle's unknowns;
    boxes thinner than min_cells dropped.  eta on the handle's nodes; origin,
    gl, gh, dom (x, y, z); shape (nz, ny, nx)."""
    tags = tag_cells(eta, dx, threshold)
    nodes = shape[::-1]
    lo = [gl[a] for a in range(3)]
    hi = [nodes[a] - 1 - gh[a] for a in range(3)]
    sub = tags[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    if sub.size == 0 or not sub.any():
        return []
    out = []
    for clo, chi in cluster_boxes(sub, tuple(origin[a] + lo[a] for a in range(3)), buffer, blocking_factor, grid_eff):
        c0, c1 = [], []
        for a in range(3):
            ulo, uhi = origin[a] + gl[a], origin[a] + nodes[a] - 1 - gh[a]
            c0.append(max(clo[a], ulo + (1 if ulo > 0 else 0)))
            c1.append(min(chi[a], uhi - (1 if uhi < dom[a] else 0)))
        if all(c1[a] - c0[a] >= min_cells for a in range(3)):
            out.append((tuple(c0), tuple(c1)))
    return out


def hierarchy(n0, eta_of, levels=3, threshold=0.1, buffer=2, blocking_factor=0, grid_eff=0.7):
    """Patches of a base of n0^3 cells (unit spacing n0 per axis of the domain)
    level by level from the oracle's tags: [(parent index, clo, chi,
    origin, gl, gh, n_cell, dom)], parents before children.  eta_of(origin,
    shape, h) gives the indicator on a box's nodes at spacing h = 1 / dom."""
    out = []
    front = [(-1, (0, 0, 0), (0, 0, 0), (0, 0, 0), (n0,) * 3, (n0,) * 3)]  # (index, origin, gl, gh, n_cell, dom)
    for _ in range(levels - 1):
        nxt = []
        for idx, origin, gl, gh, nc, dom in front:
            shape = tuple(nc[a] + 1 for a in (2, 1, 0))
            h = 1.0 / dom[0]
            boxes = child_boxes(eta_of(origin, shape, h), origin, shape, gl, gh, dom, (h,) * 3, threshold, buffer,
                                blocking_factor, grid_eff)
            for clo, chi in boxes:
                o, g0, g1, ncc = patch_geometry(dom, clo, chi)
                out.append((idx + 1, clo, chi, o, g0, g1, ncc, tuple(2 * d for d in dom)))
                nxt.append((len(out) - 1, o, g0, g1, ncc, tuple(2 * d for d in dom)))
        if not nxt:
            break
        front = nxt
    return out
