"""Measurements of the secondary configurations (the headline one is bench.py).

  python -m alamo_amd.measure --workload polycrystal257 [--fields 32]   257^3 polycrystals, 256 grains
  python -m alamo_amd.measure --workload amr3                            3-level AMR, 65^3 base

polycrystal257: unseen seeded Voronoi fields (256 grains, cubic C11=1.7, C12=1.2,
C44=0.75, Haar orientations, uniaxial 1e-3 along z) at 257^3 nodes solved to
1e-8 with the plain FP64 MLMG (V and F cycles), the FP32 path (FP32 cycles
inside FP64 FGMRES, elastic.fast_path) and the learned + FP32 path (the trained network on the
33^3 level, judged initial correction); times, iterations and the relative
L2 error against an FP64 solve to 1e-10.

amr3: three seeded spherical inclusions (r ~ U[0.05, 0.15], E ~
U[2, 20], dilatational eigenstrain 1e-3, interfaces two base cells wide),
zero-displacement faces, base 65^3 nodes, two patch levels from the
indicator-gradient tags (clustered into several patches per level);
composite V-cycle solve to 1e-8 in FP64 and with FP32 base cycles.

Results go to stdout as one JSON object per workload; with --fields the
polycrystal run also prints a summary over the fields (medians and ranges).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def _log(msg):
    print(f"[measure] {msg}", file=sys.stderr, flush=True)


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t


def polycrystal257(fields=4, n_cell=256, weights=None):
    import torch

    from . import bc as B
    from . import elastic as el
    from . import model as M
    from . import train
    net = None
    if weights and os.path.exists(weights):
        net, meta = train.load(weights)
    rows = []
    for f in range(fields):
        seed = 900000 + f
        C = train.polycrystal_field(n_cell, 256, seed, 1.7, 1.2, 0.75)
        op = el.ElasticOperator(n_cell, 1.0 / n_cell, M.MAJORMINOR)
        op.SetBC(B.uniaxial(2, 1e-3))
        op.SetModel(C)
        op.prepare()
        b = op.bc_init(torch.zeros((3,) + op.level_shape(0), dtype=torch.float64, device="cuda"))
        level = op.num_levels - 1
        while level > 0 and op.level_shape(level)[0] < 33:
            level -= 1
        if net is not None and op.level_shape(level) == (33, 33, 33):
            op.set_learned(net, level)
        ref = torch.zeros_like(b)
        el.Linear(op, cycle="F", outer="fgmres", abort_on_fail=False).solve(ref, b, 1e-10, 0.0)
        row = dict(field=seed, grains=256, nodes=int(np.prod(op.level_shape(0))), network_level=level)
        fast = el.fast_path(row["nodes"])
        runs = [("plain_fp64_V", dict(cycle="V")), ("plain_fp64_F", dict(cycle="F")), ("fp32", fast)]
        if net is not None:
            runs.append(("learned_fp32", dict(fast, learned="guess")))
        for name, kw in runs:
            x = torch.zeros_like(b)
            s = el.Linear(op, abort_on_fail=False, max_iter=400, **kw)
            _, dt = _timed(lambda: s.solve(x, b, 1e-8, 0.0))
            err = float(torch.linalg.norm(x - ref) / torch.linalg.norm(ref))
            row[name] = dict(seconds=dt, iters=s.info.iters, converged=s.info.status == 0,
                             learned_steps=s.info.learned_steps, rel_l2_err_vs_fp64=err)
            _log(f"field {seed}: {name}: {dt:.3f}s, {s.info.iters} iterations, err vs FP64 {err:.2e}")
        rows.append(row)
        del op, C
        torch.cuda.empty_cache()
    return dict(workload=f"polycrystal-{n_cell + 1}^3-256g uniaxial 1e-3 z, to 1e-8", fields=rows,
                summary=_summary(rows))


def _summary(rows):
    """Median, min and max over the fields of each path's time, iterations and
    error against the FP64 solve."""
    out = {}
    for name in ("plain_fp64_V", "plain_fp64_F", "fp32", "learned_fp32"):
        got = [r[name] for r in rows if name in r]
        if not got:
            continue
        out[name] = {k: dict(median=float(np.median([g[k] for g in got])), min=float(min(g[k] for g in got)),
                             max=float(max(g[k] for g in got)))
                     for k in ("seconds", "iters", "rel_l2_err_vs_fp64")}
        out[name]["converged"] = sum(bool(g["converged"]) for g in got)
        out[name]["fields"] = len(got)
    if "plain_fp64_V" in out:
        for name in ("fp32", "learned_fp32"):
            if name in out:
                out[name]["speedup_vs_plain_fp64_V_median"] = (out["plain_fp64_V"]["seconds"]["median"] /
                                                               out[name]["seconds"]["median"])
    return out


def _amr_models(ops, spheres, width):
    """Set each level's model and Newton right-hand side for the inclusions."""
    import torch

    from . import problems as P
    bs = []
    for op in ops:
        shape = op.level_shape(0)
        C, F0 = P.inclusions_on_box(op.origin, shape, op.dx[0], spheres, width=width)
        Cd = torch.from_numpy(C).cuda()
        op.SetModel(Cd)
        b0 = op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda"))
        bs.append(op.newton_rhs(Cd, torch.zeros_like(b0), b0, F0=torch.from_numpy(F0).cuda()))
    return bs


def amr3(n0=64, seed=0, shift=0.03, grid_eff=0.7, blocking_factor=8):
    """3-level AMR with patches from the device refinement tags, clustered
    into boxes (Berger-Rigoutsos with grid_eff and blocking_factor; None: one
    box per level), solved in FP64 and with FP32 base cycles; then the
    inclusions move by `shift` along x and the hierarchy is regridded (new
    tags, FillPatch of the FP64 solution) and solved again from the filled
    fields and from zero."""
    import torch

    from . import bc as B
    from . import elastic as el
    from . import model as M
    from . import problems as P
    spheres = P.random_spheres(3, seed=seed)
    width = 2.0 / n0

    def eta_of(sph):
        return lambda op: torch.from_numpy(P.sphere_indicators(op.origin, op.level_shape(0), op.dx[0], sph,
                                                               width)).cuda()

    base = el.ElasticOperator(n0, 1.0 / n0, M.ISOTROPIC)
    base.SetBC(B.Constant())
    kw = dict(levels=3, blocking_factor=blocking_factor if grid_eff is not None else 0, grid_eff=grid_eff)
    ops, dt_tag = _timed(lambda: el.amr_levels(base, eta_of(spheres), **kw))
    bs = _amr_models(ops, spheres, width)
    out = dict(workload=f"3-level AMR, base {n0 + 1}^3, 3 inclusions (seed {seed})", grid_eff=grid_eff,
               blocking_factor=kw["blocking_factor"],
               spheres=[dict(center=c, radius=r, E=E) for c, r, E in spheres],
               patches=[dict(clo=list(op.clo), chi=list(op.chi)) for op in ops[1:]],
               level_nodes=[int(np.prod(op.level_shape(0))) for op in ops], tag_and_refine_seconds=dt_tag)
    depth = [0] * len(ops)
    for q, op in enumerate(ops[1:], 1):
        depth[q] = depth[ops.index(op.parent)] + 1
    out["nodes_per_level"] = [sum(n for n, d in zip(out["level_nodes"], depth) if d == lev) for lev in range(3)]
    sols = {}
    for prec in ("fp64", "fp32"):
        xs = [torch.zeros_like(b) for b in bs]
        info, dt = _timed(lambda: el.amr_solve(ops, xs, bs, 1e-8, 0.0, max_iter=400, cycle="V", precision=prec,
                                               abort_on_fail=False))
        sols[prec] = xs
        out[prec] = dict(seconds=dt, iters=info.iters, converged=info.status == 0, res=info.res)
        _log(f"3-level AMR {prec}: {dt:.3f}s, {info.iters} composite V-cycles, converged={info.status == 0}")
    out["fp32_vs_fp64_rel_l2"] = [float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
                                  for a, b in zip(sols["fp32"], sols["fp64"])]
    # the inclusions move: regrid from new tags, fill from the old solution
    moved = [((c[0] + shift, c[1], c[2]), r, E) for c, r, E in spheres]
    (ops2, xs2), dt_regrid = _timed(lambda: el.amr_regrid(ops, sols["fp64"], eta_of(moved), **kw))
    bs2 = _amr_models(ops2, moved, width)
    reg = dict(shift=shift, patches=[dict(clo=list(op.clo), chi=list(op.chi)) for op in ops2[1:]],
               regrid_seconds=dt_regrid)
    for name, x0 in (("filled", xs2), ("zero", [torch.zeros_like(b) for b in bs2])):
        xs = [v.clone() for v in x0]
        info, dt = _timed(lambda: el.amr_solve(ops2, xs, bs2, 1e-8, 0.0, max_iter=400, cycle="V", precision="fp32",
                                               abort_on_fail=False))
        reg[name] = dict(seconds=dt, iters=info.iters, converged=info.status == 0, res=info.res, res0=info.res0,
                         rhsnorm=info.rhsnorm)
        _log(f"3-level AMR regrid, from {name}: {dt:.3f}s, {info.iters} composite V-cycles, "
             f"|r0| {info.res0:.3e}, |b| {info.rhsnorm:.3e}")
    out["regrid"] = reg
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=("polycrystal257", "amr3"), required=True)
    ap.add_argument("--fields", type=int, default=4)
    ap.add_argument("--one-box", action="store_true", help="amr3: one box per level (no clustering)")
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                                                      "unet_polycrystal129_level2.npz"))
    a = ap.parse_args(argv)
    r = (polycrystal257(a.fields, weights=a.weights) if a.workload == "polycrystal257"
         else amr3(grid_eff=None if a.one_box else 0.7))
    print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
