"""The AMReX plotfile writer (alamo_amd/plotfile.py, Integrator::WritePlotFile,
src/Integrator/Integrator.cpp:796-1050) read back by the parsers the reference's
own scripts use: scripts/testlib.py readHeader (variables, dimension, time,
levels, extent) and scripts/analyze.py (cell sizes after the ratio, domain and
step lines; the box of each level from Level_<l>/Cell_H by its regular
expression), restated here because those scripts import yt, which is absent.
The data round-trips through the FAB files, the node-to-cell average is exact
on trilinear fields, and the component names follow Set::Field::Name.
Whether yt itself reads these files is unpinned (yt is absent)."""
import os
import re
from types import SimpleNamespace as NS

import numpy as np

from alamo_amd import plotfile as PF

# the two levels of a 16^3 base with one patch over its cells (4, 2, 6) .. (12, 10, 14)
BASE = NS(origin=(0, 0, 0), ghost_lo=(0, 0, 0), ghost_hi=(0, 0, 0), dx=(1 / 16,) * 3, dom=(16, 16, 16))
PATCH = NS(origin=(7, 3, 11), ghost_lo=(1, 1, 1), ghost_hi=(1, 1, 1), dx=(1 / 32,) * 3, dom=(32, 32, 32))


def _nodes(op, n):
    """Global node coordinates (z, y, x) of a handle with n cells per axis."""
    return np.meshgrid(*[(op.origin[a] + np.arange(n[a] + 1)) * op.dx[a] for a in (2, 1, 0)], indexing="ij")


def _fields():
    out = []
    for op, n in ((BASE, (16, 16, 16)), (PATCH, (18, 18, 18))):
        z, y, x = _nodes(op, n)
        disp = np.stack([1 + 2 * x - y, 3 * z + 0.5 * x * y, x + y + z])     # trilinear
        stress = np.stack([(p + 1) * x + q * y * z for p in range(3) for q in range(3)])
        eta = np.stack([x * y, 1 - x * y])
        out.append((disp, stress, eta))
    return [("eta", [f[2] for f in out]), ("disp", [f[0] for f in out]), ("stress", [f[1] for f in out])]


def _read_header(path):
    """scripts/testlib.py readHeader, restated."""
    ret = {}
    with open(path + "/Header") as f:
        f.readline()
        nvars = int(f.readline())
        ret["vars"] = [f.readline().replace("\n", "") for _ in range(nvars)]
        ret["dim"] = int(f.readline())
        ret["time"] = float(f.readline())
        ret["amrlevels"] = int(f.readline())
        ret["geom_lo"] = [float(v) for v in f.readline().split()]
        ret["geom_hi"] = [float(v) for v in f.readline().split()]
        f.readline(), f.readline(), f.readline()          # ratios, domains, steps (analyze.py)
        ret["dx"] = [[float(d) for d in f.readline().replace(" \n", "").split(" ")]
                     for _ in range(ret["amrlevels"] + 1)]
    return ret


def _cell_h_boxes(path, lev):
    """scripts/analyze.py: the 3D box pattern over Level_<lev>/Cell_H."""
    exp = re.compile(r"\(\(([0-9]*),([0-9]*),([0-9]*)\) \(([0-9]*),([0-9]*),([0-9]*)\) \(([0-9]*),([0-9]*),([0-9]*)\)\)")
    out = []
    with open(os.path.join(path, f"Level_{lev}", "Cell_H")) as f:
        for line in f.readlines():
            res = exp.findall(line)
            if res:
                out.append(tuple(int(v) for v in res[0]))
    return out


def test_component_names_follow_set_field_name():
    assert PF.component_names("disp", 3) == ["disp_x", "disp_y", "disp_z"]
    assert PF.component_names("stress", 9)[1] == "stress_xy" and PF.component_names("stress", 9)[3] == "stress_yx"
    assert PF.component_names("eta", 2) == ["eta001", "eta002"] and PF.component_names("psi", 1) == ["psi"]


def test_node_to_cell_average_is_exact_on_trilinear_fields():
    z, y, x = np.meshgrid(*[np.arange(n) * 0.5 for n in (4, 5, 6)], indexing="ij")
    f = np.stack([1 + x * y * z - 2 * x + y, 3 * z])
    c = PF.node_to_cell(f)
    zc, yc, xc = np.meshgrid(*[(np.arange(n - 1) + 0.5) * 0.5 for n in (4, 5, 6)], indexing="ij")
    assert np.allclose(c, np.stack([1 + xc * yc * zc - 2 * xc + yc, 3 * zc]), rtol=0, atol=1e-14)


def test_plotfiles_read_back_by_the_reference_parsers(tmp_path):
    fields = _fields()
    paths = PF.write_mechanics(str(tmp_path / "out"), 1, [BASE, PATCH], fields, time=0.25)
    assert [os.path.basename(p) for p in paths] == ["00001node", "00001cell"]
    names = ["eta001", "eta002", "disp_x", "disp_y", "disp_z"] + PF.component_names("stress", 9)
    for p, cell in zip(paths, (False, True)):
        h = _read_header(p)
        assert h["vars"] == names and h["dim"] == 3 and h["time"] == 0.25 and h["amrlevels"] == 1
        assert h["geom_lo"] == [0.0] * 3 and h["geom_hi"] == [1.0] * 3
        assert h["dx"] == [[1 / 16] * 3, [1 / 32] * 3]
        # level boxes: the base domain and the patch's unknowns (cells 8..23, 4..19, 12..27 at level 1)
        t = 0 if cell else 1
        top = 1 - t
        assert _cell_h_boxes(p, 0) == [(0, 0, 0, 16 - top, 16 - top, 16 - top, t, t, t)]
        assert _cell_h_boxes(p, 1) == [(8, 4, 12, 24 - top, 20 - top, 28 - top, t, t, t)]
        rnames, time, levels = PF.read(p)
        assert rnames == names and time == 0.25
        for (lo, hi, data), op, sl in zip(levels, (BASE, PATCH), (np.s_[:, :, :, :], np.s_[:, 1:-1, 1:-1, 1:-1])):
            want = np.concatenate([f[1][0 if op is BASE else 1] for f in fields])[sl]
            if cell:
                want = PF.node_to_cell(want)
            assert np.array_equal(data, want)
        # the FAB line names a little-endian double and the box; the file holds exactly the values
        for l, (lo, hi, data) in enumerate(levels):
            fn = os.path.join(p, f"Level_{l}", "Cell_D_00000")
            with open(fn, "rb") as f:
                head = f.readline().decode()
            m = re.match(r"^FAB \(\(8, \(64 11 52 0 1 12 0 1023\)\),\(8, \(8 7 6 5 4 3 2 1\)\)\)"
                         r"\(\((\d+,\d+,\d+)\) \((\d+,\d+,\d+)\) \((\d+,\d+,\d+)\)\) (\d+)\n$", head)
            assert m and int(m.group(4)) == len(names)
            assert os.path.getsize(fn) == len(head) + 8 * data.size
        with open(os.path.join(p, "Level_1", "Cell_H")) as f:
            ch = f.read().split("\n")
        assert ch[:4] == ["1", "1", str(len(names)), "0"] and "FabOnDisk: Cell_D_00000 0" in ch
        mn = [float(v) for v in ch[ch.index(f"1,{len(names)}") + 1].rstrip(",").split(",")]
        assert np.allclose(mn, levels[1][2].reshape(len(names), -1).min(axis=1), rtol=1e-14)
