#!/usr/bin/env python3
"""Headline benchmark: time to 1e-8 relative residual of the nodal elastic
solve on a 129^3-node (n_cell = 128) Voronoi polycrystal on one MI355X, plus
the HBM roofline of the Fapply kernel.

  python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one full solve of the workload from a zero initial guess with the
learned + FP32 path: the trained network's judged initial correction
(alamo_amd/weights), then FP32 multigrid cycles inside FP64 flexible GMRES;
the residual that decides convergence is always the FP64 one.  The same
solve without the network is timed beside it (the ablation).  The plain FP64 multigrid
with the reference's settings (F-cycle every iteration; also the V-cycle) is
solved once beside it, and the distance of the fast solution to the plain
FP64 solution is reported.  Inputs are resident in HBM before any timing
starts.  The CPU baseline is the FP64 oracle on the host cores, on a bounded
sample (Fapply repetitions and ten F-cycles of the same problem).

  python bench.py --cpu-configs

times instead the other workloads on the GPU and, on bounded samples, the FP64
oracle on the host cores: the 65^3 inclusion (plain V-cycle to 1e-10), the
3-level AMR workload at a 65^3 base (composite V-cycle to 1e-8) and a 257^3
polycrystal (plain F-cycle to 1e-8), one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time to 1e-8 residual, 128³ elastic (learned/FP32 vs plain FP64); Fapply GB/s")
PEAK_HBM_GBS = 8000.0
# Algorithmic bytes per node of one Fapply in FP64: u (3x8) + C + f (3x8);
# MajorMinor C holds 21 doubles (168 B), isotropic C 2 doubles (16 B).
FAPPLY_BYTES_PER_NODE = {0: 24 + 16 + 24, 1: 24 + 168 + 24, 2: 24 + 72 + 24}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-cell", type=int, default=128)
    ap.add_argument("--grains", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--fapply-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 reference solves")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM traffic passes")
    ap.add_argument("--weights", default=os.path.join(ROOT, "alamo_amd", "weights", "unet_polycrystal129_level2.npz"))
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-configs", action="store_true",
                    help="GPU and CPU-oracle times of the inclusion, AMR and 257^3 workloads instead")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# The FP64 MajorMinor Fapply kernel of the finest level (LDS-tiled): the
# roofline figure is taken on the variant that reads every node's 21 C values
# (CM 1: the tile-blocked copy, non-temporal loads), whose HBM traffic carries
# the algorithmic bytes; the variant that reads C through the palette of the
# grains' distinct tensors (CM 2) moves far fewer bytes and is reported beside it.
FAPPLY_KERNEL = "k_stencil_lds<ael::Tile<32, 8>, double, 0, true, 1, double, false>"
FAPPLY_PALETTE_KERNEL = "k_stencil_lds<ael::Tile<32, 8>, double, 0, true, 2, double, false>"
COPY_BYTES = 1 << 28  # calibration copy: 256 MiB read and 256 MiB written
# F-cycles of the FP64 oracle timed for the CPU baseline (about 9 s on 16
# host threads at 129^3), scaled to the plain FP64 solve's cycle count
CPU_CYCLES = 10


def operators(p, C):
    """The workload's operator twice: reading C through its palette (the
    default) and through the full per-node copy (ALAMO_EL_PALETTE=0 at
    SetModel)."""
    from alamo_amd import elastic as el
    ops = []
    for palette in (True, False):
        op = el.ElasticOperator(p["n_cell"], p["dx"], p["sym"])
        op.SetBC(p["bc"])
        prev = os.environ.get("ALAMO_EL_PALETTE")
        if not palette:
            os.environ["ALAMO_EL_PALETTE"] = "0"
        try:
            op.SetModel(C)
        finally:
            if prev is None:
                os.environ.pop("ALAMO_EL_PALETTE", None)
            else:
                os.environ["ALAMO_EL_PALETTE"] = prev
        ops.append(op)
    return ops


def pmc_child(args):
    """Workload of one traffic pass under rocprofv3 --pmc: the calibration copy
    (an elementwise kernel of known size) and 3 launches of each Fapply
    variant."""
    import torch

    from alamo_amd import problems
    p = problems.polycrystal(args.n_cell, ngrains=args.grains, seed=args.seed)
    ops = operators(p, torch.from_numpy(p["C"]).cuda())
    a = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)
    torch.add(a, 1.0, out=c)
    u = torch.randn((3,) + ops[0].level_shape(0), dtype=torch.float64, device="cuda")
    f = torch.empty_like(u)
    for op in ops:
        for _ in range(3):
            op.Fapply(u, out=f)
    torch.cuda.synchronize()


def measure_traffic(args):
    """HBM bytes per Fapply launch from two rocprofv3 --pmc passes (FETCH_SIZE,
    then WRITE_SIZE), each scaled by what the counter reads for the calibration
    copy.  Returns a dict, or None when rocprofv3 is unavailable or fails."""
    import csv
    import shutil
    import subprocess
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None
    out = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="alamo_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
        cmd = ["timeout", "-s", "KILL", "120", "rocprofv3", "--pmc", ctr, "--output-format", "csv", "-d", d,
               "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--pmc-child",
               "--n-cell", str(args.n_cell), "--grains", str(args.grains), "--seed", str(args.seed)]
        log(f"traffic pass {ctr}")
        r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True, cwd=d)
        files = [os.path.join(dp, fn) for dp, _, fns in os.walk(d) for fn in fns if fn.endswith("counter_collection.csv")]
        if r.returncode != 0 or not files:
            log(f"traffic pass {ctr} failed (rc {r.returncode}): {r.stderr[-300:]}")
            return None
        fap, pal, cal = [], [], []
        for row in csv.DictReader(open(files[0])):
            if row["Counter_Name"] != ctr:
                continue
            if FAPPLY_KERNEL in row["Kernel_Name"]:
                fap.append(float(row["Counter_Value"]))
            elif FAPPLY_PALETTE_KERNEL in row["Kernel_Name"]:
                pal.append(float(row["Counter_Value"]))
            elif "elementwise" in row["Kernel_Name"]:
                cal.append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if not fap or not pal or not cal:
            log(f"traffic pass {ctr}: kernels not found")
            return None
        cal_kib = max(cal)  # the 256 MiB copy is the largest elementwise kernel of the pass
        factor = (COPY_BYTES / 1024.0) / cal_kib
        for key, vals in (("copy", fap), ("palette", pal)):
            kib = float(np.mean(vals))
            out.setdefault(key, {})[ctr] = dict(raw_kib=kib, calib_kib=cal_kib, factor=factor,
                                                bytes=kib * 1024.0 * factor)
    return out


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_configs(args):
    """GPU times of three more workloads beside the FP64 oracle's on the host
    cores, each oracle leg a bounded sample: a converged solve where it takes
    seconds, else a few cycles scaled to the GPU solve's cycle count."""
    import torch

    from alamo_amd import bc as B
    from alamo_amd import elastic as el
    from alamo_amd import model as M
    from alamo_amd import problems as P
    from alamo_amd import train
    from alamo_amd.measure import _amr_models
    from oracle import amr_oracle as AO
    from oracle import oracle as O

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t

    res = dict(cpu=_cpu_model(), cores=O.num_threads(), kind="port")

    # 65^3 inclusion, the plain FP64 V-cycle to 1e-10: the whole oracle solve
    p = P.inclusion(64)
    op = el.ElasticOperator(p["n_cell"], p["dx"], p["sym"])
    op.SetBC(p["bc"])
    C = torch.from_numpy(p["C"]).cuda()
    op.SetModel(C)
    b = op.bc_init(torch.zeros((3,) + op.level_shape(0), dtype=torch.float64, device="cuda"))
    rhs = op.newton_rhs(C, torch.zeros_like(b), b, F0=torch.from_numpy(p["F0"]).cuda())
    op.prepare()
    gpu_t, gpu_it = [], 0
    for _ in range(3):
        x = torch.zeros_like(rhs)
        s = el.Linear(op, cycle="V")
        _, dt = timed(lambda: s.solve(x, rhs, 1e-10, 0.0))
        gpu_t.append(dt)
        gpu_it = s.info.iters
    shape = (p["n_cell"] + 1,) * 3
    rhs_o = O.newton_rhs(np.zeros((3,) + shape), O.bc_init(shape, p["bc"]), p["C"], p["sym"], p["bc"], p["dx"],
                         F0=p["F0"])
    t = time.perf_counter()
    _, info = O.mlmg_solve(rhs_o, p["C"], p["sym"], p["bc"], p["dx"], tol_rel=1e-10, cycle="V")
    cpu_s = time.perf_counter() - t
    res["inclusion65_fp64_V_1e-10"] = dict(gpu_s=float(np.median(gpu_t)), gpu_iters=gpu_it, cpu_s=cpu_s,
                                            cpu_iters=info["iters"], sample="the whole oracle solve")
    log(f"inclusion 65^3: GPU {np.median(gpu_t):.4f} s ({gpu_it} V-cycles), oracle {cpu_s:.2f} s "
        f"({info['iters']})")
    del op, C, rhs, b

    # 3-level AMR, 65^3 base, clustered patches: 3 oracle composite V-cycles
    n0, width = 64, 2.0 / 64
    spheres = P.random_spheres(3, seed=0)
    base = el.ElasticOperator(n0, 1.0 / n0, M.ISOTROPIC)
    base.SetBC(B.Constant())
    eta_of = (lambda op_: torch.from_numpy(P.sphere_indicators(op_.origin, op_.level_shape(0), op_.dx[0], spheres,
                                                                 width)).cuda())
    ops = el.amr_levels(base, eta_of, levels=3, blocking_factor=8, grid_eff=0.7)
    bs = _amr_models(ops, spheres, width)
    gpu = {}
    for prec in ("fp64", "fp32"):
        ts = []
        for _ in range(3):
            xs = [torch.zeros_like(v) for v in bs]
            inf_, dt = timed(lambda: el.amr_solve(ops, xs, bs, 1e-8, 0.0, max_iter=400, cycle="V", precision=prec))
            ts.append(dt)
        gpu[prec] = (float(np.median(ts)), inf_.iters)
    lv = []
    for q, o in enumerate(ops):
        Cq, _ = P.inclusions_on_box(o.origin, o.level_shape(0), o.dx[0], spheres, width=width)
        if q == 0:
            lv.append(AO.Level(Cq, M.ISOTROPIC, B.Constant(), o.dx[0], dom=o.dom))
        else:
            L = AO.Level(Cq, M.ISOTROPIC, AO.patch_bc(B.Constant(), o.ghost_lo, o.ghost_hi), o.dx[0],
                         origin=o.origin, clo=o.clo, chi=o.chi, gl=o.ghost_lo, gh=o.ghost_hi, dom=o.dom)
            L.parent = next(i for i, oo in enumerate(ops) if oo is o.parent)
            lv.append(L)
    bh = [v.cpu().numpy() for v in bs]
    ncyc = 3
    t = time.perf_counter()
    AO.amr_solve(lv, [np.zeros_like(v) for v in bh], bh, 1e-300, max_iter=ncyc, cycle="V")
    per = (time.perf_counter() - t) / ncyc
    res["amr3_65_base_V_1e-8"] = dict(
        gpu_fp64_s=gpu["fp64"][0], gpu_fp32_s=gpu["fp32"][0], gpu_iters=gpu["fp64"][1],
        nodes=[int(np.prod(o.level_shape(0))) for o in ops], cpu_s=per * gpu["fp64"][1], cpu_s_per_cycle=per,
        sample=f"{ncyc} oracle composite V-cycles (Python over the C++ level kernels), scaled to the GPU's "
               f"{gpu['fp64'][1]}")
    log(f"AMR 65^3 base: GPU {gpu['fp64'][0]:.4f} s FP64 / {gpu['fp32'][0]:.4f} s FP32 ({gpu['fp64'][1]} cycles), "
        f"oracle {per:.2f} s a cycle")
    del ops, bs, lv, bh, base

    # 257^3 polycrystal (256 grains, the first unseen field): 2 oracle F-cycles
    n = 256
    Cg = train.polycrystal_field(n, 256, 900000, 1.7, 1.2, 0.75)
    bc = B.uniaxial(2, 1e-3)
    op = el.ElasticOperator(n, 1.0 / n, M.MAJORMINOR)
    op.SetBC(bc)
    op.SetModel(Cg)
    op.prepare()
    b = op.bc_init(torch.zeros((3,) + op.level_shape(0), dtype=torch.float64, device="cuda"))
    x = torch.zeros_like(b)
    s = el.Linear(op, cycle="F")
    _, gpu_plain = timed(lambda: s.solve(x, b, 1e-8, 0.0))
    plain_it = s.info.iters
    ts = []
    for _ in range(3):
        x = torch.zeros_like(b)
        f = el.Linear(op, **el.fast_path(int(np.prod(op.level_shape(0)))))
        _, dt = timed(lambda: f.solve(x, b, 1e-8, 0.0))
        ts.append(dt)
    Ch = Cg.cpu().numpy()
    del op, Cg, b, x
    torch.cuda.empty_cache()
    shape = (n + 1,) * 3
    bo = O.bc_init(shape, bc)
    ncyc = 2
    t = time.perf_counter()
    O.mlmg_solve(bo, Ch, M.MAJORMINOR, bc, 1.0 / n, tol_rel=1e-300, max_iter=ncyc, cycle="F")
    per = (time.perf_counter() - t) / ncyc
    res["polycrystal257_F_1e-8"] = dict(gpu_plain_fp64_F_s=gpu_plain, gpu_plain_iters=plain_it,
                                        gpu_fast_path_s=float(np.median(ts)), cpu_s=per * plain_it,
                                        cpu_s_per_cycle=per,
                                        sample=f"{ncyc} oracle F-cycles scaled to the GPU plain solve's {plain_it}")
    log(f"polycrystal 257^3: GPU plain F {gpu_plain:.3f} s ({plain_it} cycles), fast path "
        f"{np.median(ts):.3f} s; oracle {per:.1f} s a cycle")
    print(json.dumps(res), flush=True)
    return 0


def main():
    args = parse()
    if args.cpu_configs:
        return cpu_configs(args)
    if args.pmc_child:
        pmc_child(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)
    import torch

    from alamo_amd import elastic as el
    from alamo_amd import problems

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    t0 = time.time()
    p = problems.polycrystal(args.n_cell, ngrains=args.grains, seed=args.seed)
    log(f"built {p['name']} on the host in {time.time() - t0:.1f}s")
    C = torch.from_numpy(p["C"]).cuda()
    op_copy = operators(p, C)[1]
    op = el.ElasticOperator(p["n_cell"], p["dx"], p["sym"])
    op.SetBC(p["bc"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.SetModel(C)  # gradient masks, compact grad-C matrices, C copies, palette
    torch.cuda.synchronize()
    set_model_s = time.perf_counter() - t0
    shape = op.level_shape(0)
    nodes = int(np.prod(shape))
    b = op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda"))
    u0 = torch.zeros_like(b)
    rhs = op.newton_rhs(C, u0, b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op.prepare()  # coarse levels and the direct bottom's inverse
    torch.cuda.synchronize()
    prepare_s = time.perf_counter() - t0
    log(f"set-up once per model: SetModel {set_model_s * 1e3:.0f} ms, prepare {prepare_s * 1e3:.0f} ms")

    # --- Fapply roofline (HIP events on the stream the kernels run on) ------
    u = torch.randn((3,) + shape, dtype=torch.float64, device="cuda")
    f = torch.empty_like(u)

    def fapply_ms_of(o, uu, ff):
        for _ in range(5):
            o.Fapply(uu, out=ff)
        stream = torch.cuda.current_stream()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.fapply_reps):
            o.Fapply(uu, out=ff)
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / args.fapply_reps

    bytes_per_launch = FAPPLY_BYTES_PER_NODE[p["sym"]] * nodes
    fapply_ms = fapply_ms_of(op_copy, u, f)          # every node's 21 C values read
    fapply_gbs = bytes_per_launch / (fapply_ms * 1e-3) / 1e9
    palette_ms = fapply_ms_of(op, u, f)              # C through the palette
    palette_gbs = bytes_per_launch / (palette_ms * 1e-3) / 1e9
    palette_rows = op.palette_size()
    assert op_copy.palette_size() == 0
    log(f"Fapply (C copy) {fapply_ms * 1e3:.1f} us/launch, {fapply_gbs:.0f} GB/s algorithmic; "
        f"with the C palette of {palette_rows} rows {palette_ms * 1e3:.1f} us")
    del op_copy
    # isotropic C (2 values per node) at the same size: the spherical inclusion
    pi = problems.inclusion(args.n_cell)
    op_iso = el.ElasticOperator(pi["n_cell"], pi["dx"], pi["sym"])
    op_iso.SetBC(pi["bc"])
    op_iso.SetModel(torch.from_numpy(pi["C"]).cuda())
    iso_ms = fapply_ms_of(op_iso, u, f)
    iso_bytes = FAPPLY_BYTES_PER_NODE[pi["sym"]] * nodes
    log(f"Fapply isotropic {iso_ms * 1e3:.1f} us/launch, {iso_bytes / (iso_ms * 1e-3) / 1e9:.0f} GB/s algorithmic")
    del op_iso

    # --- plain FP64 multigrid with the reference settings --------------------
    plain = {}
    x64 = None
    if not args.no_plain:
        for cyc in ("F", "V"):
            x = torch.zeros_like(rhs)
            s = el.Linear(op, cycle=cyc, abort_on_fail=False, max_iter=200)
            torch.cuda.synchronize()
            t = time.perf_counter()
            s.solve(x, rhs, args.tol, 0.0)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            plain[cyc] = dict(seconds=dt, iters=s.info.iters, converged=s.info.status == 0,
                              rel_res=s.info.res / max(s.info.rhsnorm, s.info.res0))
            log(f"plain FP64 {cyc}-cycle: {dt:.3f}s, {s.info.iters} iters, converged={s.info.status == 0}")
            if cyc == "F" and s.info.status == 0:
                x64 = x

    # --- learned + FP32 path, and FP32 alone -----------------------------------
    learned = ""
    if os.path.exists(args.weights) and args.n_cell == 128:
        from alamo_amd import train
        net, meta = train.load(args.weights)
        op.set_learned(net, meta["level"])
        learned = "guess"
    else:
        log("no trained network for this grid: the learned step is off")

    fast = el.fast_path(nodes)

    def fast_solve(mode):
        x = torch.zeros_like(rhs)
        s = el.Linear(op, learned=mode, **fast)
        s.solve(x, rhs, args.tol, 0.0)
        return x, s

    def timed(mode):
        for _ in range(args.warmup):
            fast_solve(mode)
        torch.cuda.synchronize()
        times = []
        for _ in range(args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            x, s = fast_solve(mode)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        err = float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64)) if x64 is not None else None
        return dict(seconds=float(np.mean(times)), fgmres_steps=s.info.iters, learned_steps=s.info.learned_steps,
                    rel_res=s.info.res / max(s.info.rhsnorm, s.info.res0), rel_l2_err_vs_fp64=err)

    fp32_only = timed("")
    log(f"FP32 path: {fp32_only['seconds']:.3f}s per solve, {fp32_only['fgmres_steps']} FGMRES steps")
    head = timed(learned) if learned else fp32_only
    fast_t = head["seconds"]
    log(f"learned + FP32 path: {fast_t:.3f}s per solve, {head['fgmres_steps']} FGMRES steps, "
        f"{head['learned_steps']} learned steps taken, rel res {head['rel_res']:.2e}")

    # --- the set-up of a model update on the same handle (one time step) ------
    resetup = {}
    for _ in range(2):  # the second, warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.SetModel(C)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        op.prepare()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        op.prepare_fp32()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        resetup = dict(set_model=t1 - t0, prepare=t2 - t1, prepare_fp32=t3 - t2, total=t3 - t0)
    log("set-up of a model update: " + ", ".join(f"{k} {v * 1e3:.1f} ms" for k, v in resetup.items()))

    traffic = None if args.no_traffic else measure_traffic(args)
    traffic_bytes = palette_traffic = None
    if traffic:
        traffic_bytes = traffic["copy"]["FETCH_SIZE"]["bytes"] + traffic["copy"]["WRITE_SIZE"]["bytes"]
        palette_traffic = traffic["palette"]["FETCH_SIZE"]["bytes"] + traffic["palette"]["WRITE_SIZE"]["bytes"]
        log(f"Fapply HBM traffic {traffic_bytes / 1e6:.1f} MB per launch "
            f"({traffic_bytes / bytes_per_launch:.2f}x the algorithmic {bytes_per_launch / 1e6:.1f} MB); "
            f"with the palette {palette_traffic / 1e6:.1f} MB")

    # --- CPU baseline: the FP64 oracle on the host cores (bounded sample) ----
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        Ch = p["C"]
        uh = u.cpu().numpy()
        O.fapply(uh, Ch, p["sym"], p["bc"], p["dx"])
        reps = []
        for _ in range(20):
            t = time.perf_counter()
            O.fapply(uh, Ch, p["sym"], p["bc"], p["dx"])
            reps.append(time.perf_counter() - t)
        cpu_fapply_s = float(np.median(reps))
        rhs_h = rhs.cpu().numpy()
        ncyc = CPU_CYCLES
        t = time.perf_counter()
        _, info = O.mlmg_solve(rhs_h, Ch, p["sym"], p["bc"], p["dx"], tol_rel=args.tol, max_iter=ncyc, cycle="F")
        cpu_cycle_s = (time.perf_counter() - t) / ncyc
        plain_iters = plain.get("F", {}).get("iters", 44)
        cpu_model = "unknown"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = dict(value=cpu_cycle_s * plain_iters, unit="s", cores=O.num_threads(), kind="port", cpu=cpu_model,
                   sample=(f"FP64 oracle: {ncyc} F-cycles of the same {p['name']} problem "
                           f"({cpu_cycle_s:.2f} s/cycle) scaled to the {plain_iters} cycles of the plain FP64 "
                           f"solve; Fapply median of 20 = {cpu_fapply_s * 1e3:.0f} ms "
                           f"({bytes_per_launch / cpu_fapply_s / 1e9:.1f} GB/s)"),
                   fapply_gbs=bytes_per_launch / cpu_fapply_s / 1e9)
        log(f"cpu baseline: {cpu['sample']}")

    out = {
        "metric": METRIC,
        "value": fast_t,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fast_t * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-in-fp64-IR" + ("+bf16-net" if learned else ""),
        "data": "synthetic",
        "config": {"workload": f"{p['name']} cubic C11=1.7 C12=1.2 C44=0.75, Haar rotations, uniaxial 1e-3 in z",
                   "n_cell": args.n_cell, "nodes": nodes, "tol_rel": args.tol,
                   "path": ("learned judged initial correction (U-Net, BF16 MFMA, on the 33^3 level) + " if learned
                            else "") + (f"FP32 cycles (nu1 = {fast['pre_smooth']}, nu2 = {fast['post_smooth']}; "
                                        f"F-cycles for the first {fast['max_fmg_iter']} steps, V-cycles after) "
                                        f"preconditioning FGMRES({fast['restart']}), FP64 residual, Krylov basis "
                                        "and direct FP64 solve on the 9^3 level")},
        "learned_fp32": head if learned else None,
        "fp32_only": fp32_only,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64_V": (plain["V"]["seconds"] / fast_t) if "V" in plain else None,
        "network_gain_vs_fp32_only": fp32_only["seconds"] / fast_t if learned else None,
        "speedup_vs_plain_fp64_F": (plain["F"]["seconds"] / fast_t) if "F" in plain else None,
        "setup_once_per_model_s": dict(resetup, first_model_on_a_new_handle={"set_model": set_model_s,
                                                                              "prepare": prepare_s}),
        "fapply": {"us_per_launch": fapply_ms * 1e3, "gbs": fapply_gbs, "bytes_per_node": FAPPLY_BYTES_PER_NODE[p["sym"]],
                   "c": "per-node copy (21 values)"},
        "fapply_palette": {"us_per_launch": palette_ms * 1e3, "gbs_algorithmic": palette_gbs,
                           "c_palette_rows": palette_rows, "traffic": palette_traffic,
                           "hbm_gbs": palette_traffic / (palette_ms * 1e-3) / 1e9 if palette_traffic else None,
                           "note": ("C read as a 2-byte row index per node into a table of the distinct tensors: "
                                    "the 216 B/node of the algorithmic figure are not all moved, so its "
                                    "gbs_algorithmic is not an HBM rate; hbm_gbs is")},
        "fapply_isotropic": {"us_per_launch": iso_ms * 1e3, "bytes_per_node": FAPPLY_BYTES_PER_NODE[pi["sym"]],
                             "gbs": iso_bytes / (iso_ms * 1e-3) / 1e9, "workload": pi["name"]},
        "roofline": {"bound": "hbm", "achieved": fapply_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": fapply_gbs / PEAK_HBM_GBS, "traffic": traffic_bytes,
                     "kernel": "k_stencil_lds<Tile<32,8>,double,Fapply,compact grad-C,C copy>",
                     "algorithmic_bytes": bytes_per_launch,
                     "traffic_passes": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
