"""Seeded test problems of the elastic solve (numpy, host side).

  inclusion(n_cell)       spherical inclusion in an isotropic matrix:
                          E=1, nu=0.3 matrix; r=0.2 sphere with E=10, nu=0.3
                          and dilatational eigenstrain F0 = 1e-3 I; all
                          boundary regions zero displacement; Affine
                          isotropic models mixed by a sharp node indicator
                          (src/Integrator/Mechanics.H:249-253).
  polycrystal(n_cell)     Voronoi polycrystal of cubic grains (C11=1.7,
                          C12=1.2, C44=0.75), each rotated by a Haar-uniform
                          rotation, under uniaxial 1e-3 strain along z
                          (bc.uniaxial); seeds uniform in the unit cube from
                          numpy's PCG64 with the given seed.

Each returns a dict with sym, C (ncomp, nz, ny, nx), optional F0
(9, nz, ny, nx), bc (Constant), dx and n_cell.  The right-hand side is built
on the device by ElasticOperator.bc_init + newton_rhs, as
Base::Mechanics::TimeStepBegin does (src/Integrator/Base/Mechanics.H:175-206).
"""
from __future__ import annotations

import numpy as np

from . import bc as _bc
from . import model as _model


def inclusion(n_cell: int = 64, E_matrix=1.0, E_incl=10.0, nu=0.3, radius=0.2, center=(0.5, 0.5, 0.5),
              eigenstrain=1e-3):
    eta = _model.sphere_indicator(n_cell, center, radius)
    lm, mm = _model.lame_from_E_nu(E_matrix, nu)
    li, mi = _model.lame_from_E_nu(E_incl, nu)
    etas = np.stack([1.0 - eta, eta])
    C = _model.mix(etas, np.array([[lm, mm], [li, mi]]))
    F0 = _model.mix(etas, np.array([np.zeros(9), (eigenstrain * np.eye(3)).ravel()]))
    return dict(sym=_model.ISOTROPIC, C=np.ascontiguousarray(C), F0=np.ascontiguousarray(F0), bc=_bc.Constant(),
                dx=1.0 / n_cell, n_cell=n_cell, name=f"inclusion-{n_cell + 1}^3")


def polycrystal(n_cell: int = 128, ngrains: int = 64, seed: int = 0, C11=1.7, C12=1.2, C44=0.75, strain=1e-3,
                convention="reference", return_grains=False):
    rng = np.random.default_rng(seed)
    seeds = rng.random((ngrains, 3))
    R = _model.haar_rotations(ngrains, seed + 1)
    table = np.stack([_model.cubic(C11, C12, C44, R[g], convention=convention) for g in range(ngrains)])
    gid = _model.voronoi_grain_ids(n_cell, seeds)
    C = _model.field_from_table(gid, table)
    out = dict(sym=_model.MAJORMINOR, C=C, F0=None, bc=_bc.uniaxial(2, strain), dx=1.0 / n_cell, n_cell=n_cell,
               name=f"polycrystal-{n_cell + 1}^3-{ngrains}g")
    if return_grains:
        out.update(grain_ids=gid, table=table, seeds=seeds, rotations=R)
    return out


def sphere_eta(x, y, z, center, radius, width=0.0):
    """Indicator of a sphere at the points (x, y, z) (broadcastable): sharp
    (r < R) for width 0, else the diffuse 0.5 (1 - tanh((r - R) / width))."""
    r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    if width <= 0:
        return (r2 < radius * radius).astype(np.float64)
    return 0.5 * (1.0 - np.tanh((np.sqrt(r2) - radius) / width))


def inclusions_on_box(origin, shape, dx, spheres, E_matrix=1.0, nu=0.3, eigenstrain=1e-3, width=0.0):
    """Isotropic C (2, nz, ny, nx) and eigenstrain F0 (9, nz, ny, nx) at the
    nodes (origin + index) * dx of a box, for spheres [(center, radius, E)]:
    matrix E_matrix, each sphere its own E, all with Poisson ratio nu, and
    dilatational eigenstrain in the spheres; models are mixed by the sphere
    indicators (src/Integrator/Mechanics.H:249-253), sharp for width 0 as in
    inclusion(), diffuse over `width` otherwise.  The field is a function of
    position, so every AMR level samples the same material."""
    nz, ny, nx = shape
    x = ((origin[0] + np.arange(nx)) * dx)[None, None, :]
    y = ((origin[1] + np.arange(ny)) * dx)[None, :, None]
    z = ((origin[2] + np.arange(nz)) * dx)[:, None, None]
    lam, mu = _model.lame_from_E_nu(E_matrix, nu)
    C = np.empty((2, nz, ny, nx))
    C[0], C[1] = lam, mu
    eig = np.zeros(shape)
    for center, radius, E in spheres:
        eta = np.broadcast_to(sphere_eta(x, y, z, center, radius, width), shape)
        li, mi = _model.lame_from_E_nu(E, nu)
        C[0] = (1.0 - eta) * C[0] + eta * li
        C[1] = (1.0 - eta) * C[1] + eta * mi
        eig = np.maximum(eig, eta)
    F0 = np.zeros((9, nz, ny, nx))
    for a in range(3):
        F0[4 * a] = eigenstrain * eig
    return np.ascontiguousarray(C), np.ascontiguousarray(F0)


def random_spheres(n=3, seed=0, r_range=(0.05, 0.15), E_range=(2.0, 20.0)):
    """n spheres with seeded centres inside the cube (kept a radius from the
    faces), radii ~ U(r_range) and moduli ~ U(E_range)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        r = rng.uniform(*r_range)
        c = rng.uniform(r, 1.0 - r, size=3)
        out.append((tuple(float(v) for v in c), float(r), float(rng.uniform(*E_range))))
    return out


def sphere_indicators(origin, shape, dx, spheres, width=0.0):
    """The sphere indicators eta (nspheres, nz, ny, nx) at the nodes
    (origin + index) * dx of a box: the fields the refinement tags of
    src/Integrator/Mechanics.H:298-313 are computed from."""
    nz, ny, nx = shape
    x = ((origin[0] + np.arange(nx)) * dx)[None, None, :]
    y = ((origin[1] + np.arange(ny)) * dx)[None, :, None]
    z = ((origin[2] + np.arange(nz)) * dx)[:, None, None]
    return np.ascontiguousarray(np.stack([np.broadcast_to(sphere_eta(x, y, z, c, r, width), shape)
                                          for c, r, _ in spheres]))


def ellipse_eta(origin, shape, dx, prob_lo=(-8.0, -8.0, -8.0), a=(1.0, 0.75, 0.5), x0=(0.0, 0.0, 0.0), eps=0.1):
    """The two-component indicator of IC::Ellipse (src/IC/Ellipse.H, single
    ellipse given by its radii a, diffuse width eps) at the nodes
    prob_lo + (origin + index) * dx of a box: eta_0 = 1/2 - 1/2 erf((r^T A r - 1)
    / eps / |A r|) with A = diag(1/a^2) and r = x - x0, eta_1 = 1 - eta_0.  At
    r = 0 the quotient is -inf and eta_0 = 1, as in the reference."""
    from scipy.special import erf

    nz, ny, nx = shape
    r = [prob_lo[0] + (origin[0] + np.arange(nx)) * dx - x0[0],
         prob_lo[1] + (origin[1] + np.arange(ny)) * dx - x0[1],
         prob_lo[2] + (origin[2] + np.arange(nz)) * dx - x0[2]]
    X = r[0][None, None, :]
    Y = r[1][None, :, None]
    Z = r[2][:, None, None]
    A = [1.0 / (a[d] * a[d]) for d in range(3)]
    q = A[0] * X * X + A[1] * Y * Y + A[2] * Z * Z
    nrm = np.sqrt((A[0] * X) ** 2 + (A[1] * Y) ** 2 + (A[2] * Z) ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = (q - 1.0) / eps / nrm
    e0 = 0.5 - 0.5 * erf(s)
    e0 = np.broadcast_to(e0, shape)
    return np.ascontiguousarray(np.stack([e0, 1.0 - e0]))


def eshelby_models(E=210.0, nu=0.3, eigenstrain=1e-3):
    """The two Affine::Isotropic models of tests/Eshelby/input:57-64 as mixing
    tables: C rows (lambda, mu) and F0 rows (9 values, 3a+b); model 1 (eta_0,
    inside the ellipse) carries the dilatational eigenstrain, model 2 none."""
    lam, mu = _model.lame_from_E_nu(E, nu)
    Ct = np.array([[lam, mu], [lam, mu]])
    Ft = np.array([(eigenstrain * np.eye(3)).ravel(), np.zeros(9)])
    return Ct, Ft
