#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden from the reference tree.

Run in the build container only (it reads /root/reference):
    python tests/golden/make_fixtures.py
Outputs (data, not source):
  majorminor_uid.json  the 81 index tuples -> slot table of
                       Set::Matrix4<3,MajorMinor> (src/Set/Matrix4_MajorMinor.H:376-423),
                       extracted from the accessor's uid comparisons;
  solid_linear_cubic.dat, solid_affine_cubic.dat
                       the reference data of tests/Solid (thermo.dat time
                       series of disp_xhi_x and trac_xhi_x), copied verbatim;
  eshelby_displacement.csv, eshelby_sigma_normal.csv
                       the line-out data of tests/Eshelby/reference (z,
                       displacement x 1e4 and normal stress x 1000/E along
                       x = y = 0.25), copied verbatim.
"""
import json
import os
import re
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def uid_table():
    src = open(os.path.join(REF, "src/Set/Matrix4_MajorMinor.H")).read()
    body = src[src.index("class Matrix4<3,Sym::MajorMinor>"):]
    body = body[body.index("Scalar & operator () (const int i, const int j, const int k, const int l)\n    {"):]
    body = body[:body.index("Util::Abort")]
    table = {}
    for m in re.finditer(r"if \(([^)]*)\)\s*return data\[\s*(\d+)\]", body):
        for uid in re.findall(r"uid==\s*(\d+)", m.group(1)):
            table[int(uid)] = int(m.group(2))
    assert len(table) == 81, len(table)
    out = []
    for uid in range(81):
        i, j, k, l = uid % 3, (uid // 3) % 3, (uid // 9) % 3, uid // 27
        out.append({"ijkl": [i, j, k, l], "slot": table[uid]})
    return out


def main():
    with open(os.path.join(HERE, "majorminor_uid.json"), "w") as f:
        json.dump(uid_table(), f, indent=0)
    shutil.copyfile(os.path.join(REF, "tests/Solid/reference/linear-cubic.dat"),
                    os.path.join(HERE, "solid_linear_cubic.dat"))
    shutil.copyfile(os.path.join(REF, "tests/Solid/reference/affine-cubic.dat"),
                    os.path.join(HERE, "solid_affine_cubic.dat"))
    for name in ("displacement", "sigma_normal"):
        shutil.copyfile(os.path.join(REF, f"tests/Eshelby/reference/Eshelby_{name}.csv"),
                        os.path.join(HERE, f"eshelby_{name}.csv"))


if __name__ == "__main__":
    main()
