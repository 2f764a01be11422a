"""Host-side construction of stiffness fields, eigenstrains and synthetic
microstructures for the nodal elastic operator.

Field conventions (shared with the C-ABI in include/alamo_elastic.h):
  * one box per level with n = n_cell + 1 nodes per axis, node index i
    fastest, then j, then k;  arrays here have numpy shape (comp, nz, ny, nx);
  * stiffness ``sym`` tags: ``ISOTROPIC`` (2 comps: lambda, mu),
    ``MAJORMINOR`` (21 slots of Set::Matrix4<3,MajorMinor>), ``DIAGONAL``
    (9 comps, A[a][b] of Set::Matrix4<3,Diagonal>).

Restated behaviour (reference paths relative to the reference root):
  * E, nu -> lambda, mu: src/Model/Solid/Linear/Isotropic.H:116-187 and
    src/Model/Solid/Affine/Isotropic.H:120-125;
  * cubic rotation: src/Set/Matrix4_MajorMinor.H:518-556 (see ``cubic``);
  * model mixing sum_n eta_n * model_n: src/Integrator/Mechanics.H:249-253;
  * Voronoi partition by nearest seed: src/IC/Voronoi.H:37-118 (node
    coordinates are used here; see ``voronoi_grain_ids``).
"""
from __future__ import annotations

import itertools
import math

import numpy as np

ISOTROPIC, MAJORMINOR, DIAGONAL = 0, 1, 2
NCOMP = {ISOTROPIC: 2, MAJORMINOR: 21, DIAGONAL: 9}

# Symmetric index pairs in the order used by the 21 MajorMinor slots: the
# slots are the upper triangle, row by row, of the 6x6 matrix over these
# pairs (the uid map of src/Set/Matrix4_MajorMinor.H:327-375).
_VPAIR = {(0, 0): 0, (0, 1): 1, (1, 0): 1, (0, 2): 2, (2, 0): 2,
          (1, 1): 3, (1, 2): 4, (2, 1): 4, (2, 2): 5}


def mm_slot(i: int, j: int, k: int, l: int) -> int:
    """Slot of C_ijkl among the 21 MajorMinor values."""
    p, q = _VPAIR[(i, j)], _VPAIR[(k, l)]
    if p > q:
        p, q = q, p
    return p * 6 - p * (p - 1) // 2 + (q - p)


def mm_full(slots) -> np.ndarray:
    """Expand 21 slots to the 3x3x3x3 tensor."""
    out = np.empty((3, 3, 3, 3))
    for i, j, k, l in itertools.product(range(3), repeat=4):
        out[i, j, k, l] = slots[mm_slot(i, j, k, l)]
    return out


def lame_from_E_nu(E: float, nu: float) -> tuple[float, float]:
    """(lambda, mu) from Young's modulus and Poisson ratio
    (src/Model/Solid/Affine/Isotropic.H:120-125)."""
    lam = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu)
    mu = E / 2.0 / (1.0 + nu)
    return lam, mu


def isotropic_slots(lam: float, mu: float) -> np.ndarray:
    """21 MajorMinor slots of the isotropic tensor
    C_ijkl = mu (d_ik d_jl + d_il d_jk) + lambda d_ij d_kl."""
    I = np.eye(3)
    out = np.empty(21)
    for i, j, k, l in itertools.product(range(3), repeat=4):
        out[mm_slot(i, j, k, l)] = mu * (I[i, k] * I[j, l] + I[i, l] * I[j, k]) + lam * I[i, j] * I[k, l]
    return out


def euler_rotation(phi1: float, Phi: float, phi2: float) -> np.ndarray:
    """R = X(phi2) Z(Phi) X(phi1), the Bunge-style composition of
    src/Set/Matrix4_MajorMinor.H:548-556."""
    def rx(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])

    def rz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    return rx(phi2) @ rz(Phi) @ rx(phi1)


def haar_rotations(n: int, seed: int) -> np.ndarray:
    """n rotation matrices uniform on SO(3) (Shoemake's quaternion method)."""
    rng = np.random.default_rng(seed)
    u1, u2, u3 = rng.random((3, n))
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def cubic(C11: float, C12: float, C44: float, R=None, convention: str = "reference") -> np.ndarray:
    """21 MajorMinor slots of a rotated cubic stiffness.

    ``convention="reference"`` reproduces Set::Matrix4<3,MajorMinor>::Cubic
    (src/Set/Matrix4_MajorMinor.H:518-547) value for value: the crystal-frame
    tensor holds C11 on iiii, C44 on ijij (i != j) and C12 on iikk (i != k),
    without the ijji entries, and each rotated C'_pqst is written through the
    symmetric accessor in p,q,s,t loop order, so the last of the symmetric
    variants of a slot sets it.  For rotations that are not axis
    permutations this differs from the rotated cubic tensor.
    ``convention="tensor"`` returns the rotated cubic tensor
    C12 d_ij d_kl + C44 (d_ik d_jl + d_il d_jk) + (C11-C12-2 C44) sum_a e_a^4.
    """
    R = np.eye(3) if R is None else np.asarray(R, dtype=np.float64)
    I = np.eye(3)
    if convention == "reference":
        ct = np.zeros((3, 3, 3, 3))
        for i, j, k, l in itertools.product(range(3), repeat=4):
            if i == j == k == l:
                ct[i, j, k, l] = C11
            elif i == k and j == l:
                ct[i, j, k, l] = C44
            elif i == j and k == l:
                ct[i, j, k, l] = C12
        rot = np.einsum("pi,qj,sk,tl,ijkl->pqst", R, R, R, R, ct)
        out = np.empty(21)
        for p, q, s, t in itertools.product(range(3), repeat=4):
            out[mm_slot(p, q, s, t)] = rot[p, q, s, t]
        return out
    if convention == "tensor":
        c = C12 * np.einsum("ij,kl->ijkl", I, I) + C44 * (
            np.einsum("ik,jl->ijkl", I, I) + np.einsum("il,jk->ijkl", I, I))
        for a in range(3):
            c[a, a, a, a] += C11 - C12 - 2.0 * C44
        rot = np.einsum("pi,qj,sk,tl,ijkl->pqst", R, R, R, R, c)
        out = np.empty(21)
        for p, q, s, t in itertools.product(range(3), repeat=4):
            out[mm_slot(p, q, s, t)] = rot[p, q, s, t]
        return out
    raise ValueError(f"unknown convention {convention!r}")


def node_coords(n_cell, prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 1.0)):
    """Node coordinate arrays (z, y, x broadcastable) of a box with n_cell cells."""
    n_cell = np.broadcast_to(np.asarray(n_cell), (3,))
    xs = [prob_lo[a] + (prob_hi[a] - prob_lo[a]) * np.arange(n_cell[a] + 1) / n_cell[a] for a in range(3)]
    return xs[2][:, None, None], xs[1][None, :, None], xs[0][None, None, :]


def voronoi_grain_ids(n_cell: int, seeds: np.ndarray, chunk: int = 8) -> np.ndarray:
    """Nearest-seed grain index at every node of the unit cube.

    Node positions i*dx are used; src/IC/Voronoi.H:74-76 evaluates cell-centre
    coordinates even on nodal fields, which is a reference quirk not copied
    here (the grain fields are this build's own seeded inputs)."""
    z, y, x = node_coords(n_cell)
    n = n_cell + 1
    out = np.empty((n, n, n), dtype=np.int32)
    for k0 in range(0, n, chunk):
        zz = z[k0:k0 + chunk]
        d = ((x[None] - seeds[:, 0, None, None, None]) ** 2 +
             (y[None] - seeds[:, 1, None, None, None]) ** 2 +
             (zz[None] - seeds[:, 2, None, None, None]) ** 2)
        out[k0:k0 + chunk] = np.argmin(d, axis=0)
    return out


def field_from_table(grain_ids: np.ndarray, table: np.ndarray) -> np.ndarray:
    """Per-node SoA field (ncomp, nz, ny, nx) from a per-grain table (ngrain, ncomp)."""
    return np.ascontiguousarray(np.moveaxis(table[grain_ids], -1, 0))


def mix(eta: np.ndarray, models: np.ndarray) -> np.ndarray:
    """model(node) = sum_n eta_n(node) * model_n  (src/Integrator/Mechanics.H:249-253).
    eta: (nmodels, nz, ny, nx); models: (nmodels, ncomp)."""
    return np.einsum("n...,nc->c...", eta, models)


def sphere_indicator(n_cell: int, center, radius: float) -> np.ndarray:
    """Sharp indicator r^2 < R^2 at the nodes (src/IC/Sphere.H:70-145)."""
    z, y, x = node_coords(n_cell)
    r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    return (r2 < radius * radius).astype(np.float64)
