"""Grids of the next AMR level from tags (AmrCore::MakeNewGrids with
amrex::ClusterList; src/Integrator/Integrator.cpp:23-140 reads
amr.blocking_factor, amr.grid_eff, amr.n_error_buf): the library's host
clustering (alamo_amr_cluster, called through the C-ABI without a GPU)
against the oracle's restatement of Berger-Rigoutsos, and the oracle pinned
by what the construction guarantees: every buffered tag is covered, boxes
are aligned to blocking_factor / 2 cells, no two boxes come closer than one
cell, a box that is all tags comes back whole, separated blobs come back as
their own aligned boxes, and a box cut at a hole of the signature.  The
composite solve over several patches per level reproduces the
UniaxialTension closed form (tests/UniaxialTension/test:27-28).
AMReX's own ClusterList is absent, so its exact cuts are parity unpinned."""
import ctypes

import numpy as np
import pytest

from alamo_amd import bc as B
from alamo_amd import model as M
from oracle import amr_oracle as AO
from oracle import oracle as O


def lib_boxes(tags, origin, buffer, bf, eff):
    from alamo_amd import elastic as el
    t = np.ascontiguousarray(np.asarray(tags, dtype=np.uint8))
    nz, ny, nx = t.shape
    out = (ctypes.c_int * 6000)()
    nb = ctypes.c_int()
    rc = el.lib().alamo_amr_cluster(t.ctypes.data, (ctypes.c_int * 3)(nx, ny, nz), (ctypes.c_int * 3)(*origin),
                                    buffer, bf, eff, 1000, out, ctypes.byref(nb))
    assert rc == 0, el.lib().alamo_el_last_error()
    return [(tuple(out[6 * q:6 * q + 3]), tuple(out[6 * q + 3:6 * q + 6])) for q in range(nb.value)]


def random_tags(rng):
    shape = tuple(int(v) for v in rng.integers(3, 30, size=3))
    k, j, i = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    t = np.zeros(shape, bool)
    for _ in range(int(rng.integers(1, 6))):
        c = rng.integers(0, shape)
        r = int(rng.integers(1, 6))
        t |= (k - c[0]) ** 2 + (j - c[1]) ** 2 + (i - c[2]) ** 2 < r * r
    t |= rng.random(shape) < 0.01 * rng.random()
    return t


def test_library_clustering_matches_oracle():
    rng = np.random.default_rng(0)
    for _ in range(150):
        t = random_tags(rng)
        origin = tuple(int(v) for v in rng.integers(0, 20, size=3))
        buf, bf = int(rng.integers(0, 3)), int(rng.choice([1, 2, 4, 8]))
        eff = float(rng.choice([0.5, 0.7, 0.9, 1.0]))
        assert lib_boxes(t, origin, buf, bf, eff) == AO.cluster_boxes(t, origin, buf, bf, eff)


def test_oracle_clustering_guarantees():
    rng = np.random.default_rng(1)
    for _ in range(100):
        t = random_tags(rng)
        origin = tuple(int(v) for v in rng.integers(0, 20, size=3))
        buf, bf = int(rng.integers(0, 3)), int(rng.choice([2, 4, 8]))
        eff = float(rng.choice([0.5, 0.8, 1.0]))
        boxes = AO.cluster_boxes(t, origin, buf, bf, eff)
        nz, ny, nx = t.shape
        grown = t.copy()
        k, j, i = np.nonzero(t)
        for dk in range(-buf, buf + 1):
            for dj in range(-buf, buf + 1):
                for di in range(-buf, buf + 1):
                    grown[np.clip(k + dk, 0, nz - 1), np.clip(j + dj, 0, ny - 1), np.clip(i + di, 0, nx - 1)] = True
        cover = np.zeros_like(t)
        q = bf // 2
        n = (nx, ny, nz)
        for clo, chi in boxes:
            for a in range(3):
                assert clo[a] % q == 0 or clo[a] == origin[a]            # aligned, or at the array's edge
                assert chi[a] % q == 0 or chi[a] == origin[a] + n[a]
            cover[clo[2] - origin[2]:chi[2] - origin[2], clo[1] - origin[1]:chi[1] - origin[1],
                  clo[0] - origin[0]:chi[0] - origin[0]] = True
        assert not (grown & ~cover).any()                                # every buffered tag covered
        for x in range(len(boxes)):
            for y in range(x + 1, len(boxes)):
                (a0, a1), (b0, b1) = boxes[x], boxes[y]
                assert any(a1[a] < b0[a] or b1[a] < a0[a] for a in range(3))   # at least one cell apart


def test_oracle_clustering_known_answers():
    t = np.zeros((16, 16, 16), bool)
    t[2:6, 3:7, 1:5] = True                                   # a full block: one box, itself
    assert AO.cluster_boxes(t, (0, 0, 0), 0, 1, 1.0) == [((1, 3, 2), (5, 7, 6))]
    t[10:14, 9:12, 11:15] = True                              # a second blob far away: its own box
    assert AO.cluster_boxes(t, (0, 0, 0), 0, 1, 1.0) == [((1, 3, 2), (5, 7, 6)), ((11, 9, 10), (15, 12, 14))]
    # aligned to blocking_factor / 2 = 4 cells on the global grid (origin 2)
    assert AO.cluster_boxes(t, (2, 2, 2), 0, 8, 1.0) == [((2, 4, 4), (8, 12, 8)), ((12, 8, 12), (18, 16, 16))]
    # an L: two bars meeting at a corner; Berger-Rigoutsos cuts it (no hole:
    # at the inflection of the signature) into boxes of tags only, and the
    # touching pieces are then merged back into one box for the solver
    t = np.zeros((4, 12, 12), bool)
    t[:, 0:2, :] = True
    t[:, :, 0:2] = True
    raw = []
    AO._berger_rigoutsos(t, ((0, 0, 0), (11, 11, 3)), 1.0, raw)
    assert len(raw) >= 2
    cells = sum((h[0] - l[0] + 1) * (h[1] - l[1] + 1) * (h[2] - l[2] + 1) for l, h in raw)
    assert cells == int(t.sum())                             # efficiency 1: exactly the tags
    assert AO.cluster_boxes(t, (0, 0, 0), 0, 1, 1.0) == [((0, 0, 0), (12, 12, 4))]
    # a hole: two slabs with an empty plane between are cut there and, one
    # cell apart, stay two boxes
    t = np.zeros((6, 5, 9), bool)
    t[:, :, 0:4] = True
    t[:, :, 5:9] = True
    assert AO.cluster_boxes(t, (0, 0, 0), 0, 1, 1.0) == [((0, 0, 0), (4, 5, 6)), ((5, 0, 0), (9, 5, 6))]


def test_uniaxial_tension_closed_form_on_several_patches_per_level():
    """Two separated patches on level 1 and one under each on level 2 (a tree
    of five handles): the composite solution is the closed form on all of
    them."""
    from test_oracle_closed_forms import uniaxial_tension_deck_bc
    n0 = 16
    bc = uniaxial_tension_deck_bc()
    lam, mu = M.lame_from_E_nu(210.0, 0.3)
    spec = [(0, (2, 3, 3), (6, 12, 12)), (0, (9, 4, 5), (14, 11, 12)),    # level 1, parent base
            (1, (6, 8, 8), (10, 20, 20)), (2, (20, 10, 12), (26, 18, 22))]  # level 2, parents 1 and 2
    dom0 = (n0,) * 3
    h0 = 16.0 / n0
    levels = [AO.Level(np.stack([np.full((n0 + 1,) * 3, lam), np.full((n0 + 1,) * 3, mu)]), M.ISOTROPIC, bc, h0,
                       dom=dom0)]
    for par, clo, chi in spec:
        pdom = levels[par].dom
        origin, gl, gh, nc = AO.patch_geometry(pdom, clo, chi)
        shape = tuple(nc[a] + 1 for a in (2, 1, 0))
        lv = AO.Level(np.stack([np.full(shape, lam), np.full(shape, mu)]), M.ISOTROPIC, AO.patch_bc(bc, gl, gh),
                      levels[par].dx / 2, origin=origin, clo=clo, chi=chi, gl=gl, gh=gh,
                      dom=tuple(2 * d for d in pdom))
        lv.parent = par
        levels.append(lv)
    bs = [O.bc_init(lv.shape, lv.bc) for lv in levels]
    xs, info = AO.amr_solve(levels, [np.zeros_like(b) for b in bs], bs, 1e-10, max_iter=200)
    assert info["status"] == 0
    eps = 0.1 / 16
    for lv, x in zip(levels, xs):
        nx = lv.shape[2]
        ux = (np.arange(nx) + lv.origin[0]) * lv.dx * eps
        assert np.abs(x[0] - ux[None, None, :]).max() < 1e-9 * 0.1
