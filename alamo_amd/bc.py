"""Constant elastic boundary conditions on the 26 boundary regions of a box.

Mirrors BC::Operator::Elastic::Constant (src/BC/Operator/Elastic/Constant.H):
every region (6 faces, 12 edges, 8 corners) carries, per displacement
component, a type (displacement, traction, periodic, neumann, none) and a
value.  The default is zero displacement everywhere (Constant.H:36-45).
Region order and names follow BC::Operator::Elastic::Elastic::Face
(src/BC/Operator/Elastic/Elastic.H:35-43) and the input keys of
Constant::Parse (Constant.H:439-648).
"""
from __future__ import annotations

import numpy as np

DISPLACEMENT, TRACTION, PERIODIC, NEUMANN, NONE = 0, 1, 2, 3, 4

FACES = [
    "xlo", "ylo", "zlo", "xhi", "yhi", "zhi",
    "ylozlo", "ylozhi", "yhizlo", "yhizhi",
    "zloxlo", "zloxhi", "zhixlo", "zhixhi",
    "xloylo", "xloyhi", "xhiylo", "xhiyhi",
    "xloylozlo", "xloylozhi", "xloyhizlo", "xloyhizhi",
    "xhiylozlo", "xhiylozhi", "xhiyhizlo", "xhiyhizhi",
]
FACE_INDEX = {name: i for i, name in enumerate(FACES)}

_TYPE_NAMES = {"displacement": DISPLACEMENT, "disp": DISPLACEMENT, "neumann": NEUMANN,
               "traction": TRACTION, "trac": TRACTION, "periodic": PERIODIC}


class Constant:
    """Per-region, per-component boundary type and value table."""

    def __init__(self):
        self.type = np.full((26, 3), DISPLACEMENT, dtype=np.int32)
        self.val = np.zeros((26, 3), dtype=np.float64)

    def set(self, face, direction: int, type_: int, value: float = 0.0) -> "Constant":
        f = FACE_INDEX[face] if isinstance(face, str) else int(face)
        self.type[f, direction] = type_
        self.val[f, direction] = value
        return self

    @classmethod
    def parse(cls, entries: dict) -> "Constant":
        """Build from reference-style keys, e.g. ``{"type.xlo": "disp trac trac",
        "val.xhi": "0.1 0.0 0.0"}`` (Constant::Parse).  Unknown type strings
        raise, as the reference's map lookup would yield an invalid type."""
        bc = cls()
        for key, value in entries.items():
            kind, _, face = key.partition(".")
            words = value.split() if isinstance(value, str) else list(value)
            if len(words) != 3:
                raise ValueError(f"{key}: expected 3 entries, got {value!r}")
            f = FACE_INDEX[face]
            if kind == "type":
                bc.type[f] = [_TYPE_NAMES[w] for w in words]
            elif kind == "val":
                bc.val[f] = [float(w) for w in words]
            else:
                raise KeyError(key)
        return bc

    def regions_touching(self, face: str):
        """Indices of the regions (face, edges, corners) lying on one face."""
        return [i for i, name in enumerate(FACES) if face in _split(name)]


def _split(name: str):
    return [name[i:i + 3] for i in range(0, len(name), 3)]


def uniaxial(axis: int, strain: float, length: float = 1.0) -> Constant:
    """Uniaxial stress along ``axis`` on a box: the lo face of ``axis`` holds
    u_axis = 0, the hi face u_axis = strain*length, the lo faces of the two
    other axes are rollers (zero normal displacement) and every other
    component is traction free.  Edges and corners take, per component, the
    displacement condition of any face they lie on and are traction free
    otherwise (the pattern of tests/UniaxialTension/input, along any axis)."""
    names = "xyz"
    bc = Constant()
    bc.type[:] = TRACTION
    bc.val[:] = 0.0
    for r, name in enumerate(FACES):
        parts = _split(name)
        for c in range(3):
            if c == axis:
                if names[axis] + "lo" in parts:
                    bc.type[r, c], bc.val[r, c] = DISPLACEMENT, 0.0
                elif names[axis] + "hi" in parts:
                    bc.type[r, c], bc.val[r, c] = DISPLACEMENT, strain * length
            elif names[c] + "lo" in parts:
                bc.type[r, c], bc.val[r, c] = DISPLACEMENT, 0.0
    return bc


def shear(normal: int, direction: int, strain: float, length: float = 1.0) -> Constant:
    """Simple shear of a box: every region on the lo face of ``normal`` is
    clamped, every region on its hi face moves by strain*length along
    ``direction`` with the other components held at 0, and the remaining
    faces are traction free."""
    if normal == direction:
        raise ValueError("shear needs direction != normal")
    names = "xyz"
    bc = Constant()
    bc.type[:] = TRACTION
    bc.val[:] = 0.0
    for r, name in enumerate(FACES):
        parts = _split(name)
        if names[normal] + "lo" in parts:
            bc.type[r, :] = DISPLACEMENT
        elif names[normal] + "hi" in parts:
            bc.type[r, :] = DISPLACEMENT
            bc.val[r, direction] = strain * length
    return bc


def tension_test(type_: str = "uniaxial_stress_clamp", disp: float | None = None, trac: float | None = None) -> Constant:
    """BC::Operator::Elastic::TensionTest::Parse (src/BC/Operator/Elastic/TensionTest.H:48-306)
    for one load value: tension along x with ``disp`` (displacement of the
    x-hi side) or ``trac`` (traction on it); exactly one must be given."""
    if (disp is None) == (trac is None):
        raise ValueError("specify exactly one of disp and trac")
    mode_disp = disp is not None
    load = disp if mode_disp else trac
    # the clamp, kolsky and strain variants write the displacement input
    # (default "0.0") even in traction mode
    dval = disp if mode_disp else 0.0
    bc = Constant()
    F = FACE_INDEX
    xhi_regions = ["xhi", "xhiylo", "xhiyhi", "zloxhi", "zhixhi",
                   "xhiylozlo", "xhiylozhi", "xhiyhizlo", "xhiyhizhi"]
    xlo_regions = ["xlo", "xloylo", "xloyhi", "zloxlo", "zhixlo",
                   "xloylozlo", "xloylozhi", "xloyhizlo", "xloyhizhi"]
    if type_ == "uniaxial_stress_clamp":
        for r in xhi_regions:
            bc.val[F[r], 0] = dval
        for r in ["ylo", "yhi", "zlo", "zhi", "ylozlo", "yhizlo", "ylozhi", "yhizhi"]:
            bc.type[F[r], :] = TRACTION
    elif type_ == "uniaxial_kolsky":
        for r in ["xlo", "xloylo", "xloyhi"]:
            bc.val[F[r], 0] = dval
        for r in ["ylo", "yhi"]:
            bc.type[F[r], :] = TRACTION
    elif type_ in ("uniaxial_stress", "uniaxial_strain"):
        bc.type[:] = TRACTION
        for r in xlo_regions:
            bc.type[F[r], 0] = DISPLACEMENT
        if type_ == "uniaxial_strain" or mode_disp:
            for r in xhi_regions:
                bc.type[F[r], 0], bc.val[F[r], 0] = DISPLACEMENT, dval
        else:
            for r in xhi_regions:
                bc.type[F[r], 0], bc.val[F[r], 0] = TRACTION, load
        if type_ == "uniaxial_stress":
            for r in ["ylo", "xloylo", "xhiylo", "ylozlo", "ylozhi", "xloylozlo", "xloylozhi", "xhiylozlo",
                      "xhiylozhi"]:
                bc.type[F[r], 1] = DISPLACEMENT
            for r in ["zlo", "ylozlo", "yhizlo", "zloxlo", "zloxhi", "xloylozlo", "xloyhizlo", "xhiylozlo",
                      "xhiyhizlo"]:
                bc.type[F[r], 2] = DISPLACEMENT
        else:
            for r in ["ylo", "xloylo", "xhiylo", "yhi", "xloyhi", "xhiyhi", "ylozlo", "ylozhi", "yhizlo", "yhizhi",
                      "xloylozlo", "xloylozhi", "xhiylozlo", "xhiylozhi", "xloyhizlo", "xloyhizhi", "xhiyhizlo",
                      "xhiyhizhi"]:
                bc.type[F[r], 1] = DISPLACEMENT
            for r in ["zlo", "zhi", "ylozlo", "yhizlo", "zloxlo", "zloxhi", "ylozhi", "yhizhi", "zhixlo", "zhixhi",
                      "xloylozlo", "xloyhizlo", "xhiylozlo", "xhiyhizlo", "xloylozhi", "xloyhizhi", "xhiylozhi",
                      "xhiyhizhi"]:
                bc.type[F[r], 2] = DISPLACEMENT
    else:
        raise ValueError(f"Unknown BC type {type_}")
    return bc
