"""Training of the learned coarse-level solver on solutions that the solver
itself generates.

  python -m alamo_amd.train --samples 1024 --val 32 --epochs 80 --out weights.npz

Each sample is a seeded Voronoi polycrystal on the finest grid (16-128
grains, cubic C11 ~ U[1, 5] with C12 and C44 in the ratios of C11 = 1.7,
C12 = 1.2, C44 = 0.75, Haar-random orientation per grain) under a random
load of magnitude 1e-3: uniaxial tension along a random axis, simple shear
on a random face (and, with --body-fraction, a random smooth body force with
all faces clamped).  The operator is averaged down to MG level L;
the level-L right-hand side (alamo_el_coarse_rhs) is solved there exactly in
FP64 (FGMRES with an F-cycle, to 1e-10) and the pair (network input, u
scaled to unit RMS) is kept on the GPU: the solver uses the output as a
direction and picks the step length itself.  The network is trained with the library's own forward,
backward and Adam kernels (alamo_nn_train_step), one sample per step, with
a cosine learning-rate schedule.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import bc as _bc
from . import elastic as el
from . import model as _model
from .learned import UNet

C_RATIOS = (1.0, 1.2 / 1.7, 0.75 / 1.7)
LOAD = 1e-3


def _log(msg):
    print(f"[train] {msg}", flush=True)


def random_load(rng, body_fraction=0.0):
    """(kind, Constant table, body force flag) of one sample: uniaxial tension
    and simple shear with equal odds, body forces with odds body_fraction."""
    r = rng.random()
    sign = 1.0 if rng.random() < 0.5 else -1.0
    if r < 0.5 * (1.0 - body_fraction):
        axis = int(rng.integers(3))
        return f"uniaxial-{'xyz'[axis]}", _bc.uniaxial(axis, sign * LOAD), False
    if r < 1.0 - body_fraction:
        normal = int(rng.integers(3))
        direction = int((normal + 1 + rng.integers(2)) % 3)
        return f"shear-{'xyz'[normal]}{'xyz'[direction]}", _bc.shear(normal, direction, sign * LOAD), False
    return "body-force", _bc.Constant(), True


def smooth_force(n_cell, rng, modes=4):
    """Random smooth body force on the nodes (3, n, n, n), zero on the boundary rows."""
    x = np.linspace(0.0, 1.0, n_cell + 1)
    f = np.zeros((3, n_cell + 1, n_cell + 1, n_cell + 1))
    for c in range(3):
        for _ in range(modes):
            kx, ky, kz = rng.integers(1, 5, size=3)
            a = rng.standard_normal()
            ph = rng.random(3) * np.pi
            f[c] += a * (np.sin(np.pi * kz * x + ph[2])[:, None, None] * np.sin(np.pi * ky * x + ph[1])[None, :, None]
                         * np.sin(np.pi * kx * x + ph[0])[None, None, :])
    f[:, [0, -1], :, :] = 0.0
    f[:, :, [0, -1], :] = 0.0
    f[:, :, :, [0, -1]] = 0.0
    return f


def polycrystal_field(n_cell, ngrains, seed, C11, C12, C44):
    """Per-node MajorMinor stiffness (21, n, n, n) float64 on the GPU: the
    field of problems.polycrystal (same seeds, rotations and nearest-seed
    rule at the node positions i*dx) built with torch on the device, so that
    a training sample costs milliseconds instead of host seconds."""
    import torch
    rng = np.random.default_rng(seed)
    seeds = torch.from_numpy(rng.random((ngrains, 3))).cuda()
    R = _model.haar_rotations(ngrains, seed + 1)
    table = torch.from_numpy(np.stack([_model.cubic(C11, C12, C44, R[g]) for g in range(ngrains)])).cuda()
    n = n_cell + 1
    xs = torch.arange(n, dtype=torch.float64, device="cuda") / n_cell
    gid = torch.empty((n, n, n), dtype=torch.long, device="cuda")
    dxy = (xs[None, None, :] - seeds[:, 0, None, None]) ** 2 + (xs[None, :, None] - seeds[:, 1, None, None]) ** 2
    for k in range(n):
        d = dxy + ((xs[k] - seeds[:, 2]) ** 2)[:, None, None]
        gid[k] = torch.argmin(d, dim=0)
    return table[gid].permute(3, 0, 1, 2).contiguous()


def make_sample(k, n_cell, level, seed0=1000, tol=1e-10, ngrains=None, load=None, body_fraction=0.0):
    """One (input, target) pair on MG level `level` of an n_cell^3 problem.
    Returns a dict with float32 CUDA tensors x (n_in, ...), y (3, ...)."""
    import torch
    rng = np.random.default_rng(seed0 + k)
    ng = int(rng.integers(16, 129)) if ngrains is None else int(ngrains)
    C11 = float(rng.uniform(1.0, 5.0))
    kind, bct, body = random_load(rng, body_fraction) if load is None else load
    C = polycrystal_field(n_cell, ng, seed0 + k, C11, C11 * C_RATIOS[1], C11 * C_RATIOS[2])
    op = el.ElasticOperator(n_cell, 1.0 / n_cell, _model.MAJORMINOR)
    op.SetBC(bct)
    op.SetModel(C)
    op.prepare()
    shape = op.level_shape(0)
    if body:
        b = torch.from_numpy(smooth_force(n_cell, rng) * C11).cuda()
    else:
        b = op.bc_init(torch.zeros((3,) + shape, dtype=torch.float64, device="cuda"))
    bc = op.coarse_rhs(b, level)
    x, sc = op.learned_input(bc, level)
    nc = n_cell >> level
    opc = el.ElasticOperator(nc, float(1 << level) / n_cell, _model.MAJORMINOR)
    opc.SetBC(bct)
    opc.SetModel(op.level_model(level))
    u = torch.zeros_like(bc)
    s = el.Linear(opc, cycle="F", outer="fgmres", restart=30, abort_on_fail=False, max_iter=300)
    s.solve(u, bc, tol, 0.0)
    ok = s.info.status == 0
    # The solver takes the network output as a direction (the step length is
    # chosen by residual minimization), so targets are scaled to unit RMS:
    # boundary and body-force samples weigh alike in the loss.
    rms = float(torch.sqrt(torch.mean(u * u)))
    y = (u / rms).float() if rms > 0 else torch.zeros_like(u, dtype=torch.float32)
    return dict(x=x, y=y, scale=sc, kind=kind, ngrains=ng, C11=C11, converged=ok, iters=s.info.iters)


def generate(n, n_cell, level, seed0=1000, body_fraction=0.0):
    data = []
    t0 = time.time()
    for k in range(n):
        d = make_sample(k, n_cell, level, seed0=seed0, body_fraction=body_fraction)
        if not d["converged"]:
            _log(f"sample {k}: coarse solve did not converge, skipped")
            continue
        data.append(d)
        if (k + 1) % 128 == 0 or k + 1 == n:
            _log(f"generated {k + 1}/{n} samples ({time.time() - t0:.1f}s)")
    return data


def rel_errors(net, data):
    import torch
    errs = []
    for d in data:
        yp = net(d["x"])
        errs.append(float(torch.linalg.norm(yp - d["y"]) / torch.linalg.norm(d["y"]).clamp_min(1e-30)))
    return errs


def train(net, data, val, epochs, lr_max=1e-3, lr_min=1e-5, seed=0):
    rng = np.random.default_rng(seed)
    steps = epochs * len(data)
    step = 0
    hist = []
    for ep in range(epochs):
        t0 = time.time()
        losses = []
        for i in rng.permutation(len(data)):
            lr = lr_min + 0.5 * (lr_max - lr_min) * (1.0 + math.cos(math.pi * step / max(1, steps)))
            losses.append(net.train_step(data[i]["x"], data[i]["y"], lr))
            step += 1
        errs = rel_errors(net, val) if val else []
        ve = float(np.median(errs)) if errs else float("nan")
        per_kind = {}
        for d, e in zip(val, errs):
            per_kind.setdefault(d["kind"].split("-")[0], []).append(e)
        pk = {k: float(np.median(v)) for k, v in per_kind.items()}
        hist.append(dict(epoch=ep + 1, loss=float(np.mean(losses)), val_rel_err=ve, val_by_kind=pk,
                         seconds=time.time() - t0))
        _log(f"epoch {ep + 1}/{epochs}: loss {np.mean(losses):.4e}, validation median rel L2 {ve:.4f} "
             f"{ {k: round(v, 4) for k, v in pk.items()} } ({time.time() - t0:.1f}s)")
    return hist


def save(path, net, meta):
    p = net.params().cpu().numpy()
    np.savez_compressed(path, params=p, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))


def load(path, nodes=None):
    """UNet from a file written by save(); nodes overrides the grid."""
    import torch
    z = np.load(path)  # allow_pickle stays False
    meta = json.loads(bytes(z["meta"]).decode())
    net = UNet(tuple(nodes) if nodes is not None else tuple(meta["nodes"]), meta["n_in"], meta["n_out"], seed=None)
    net.set_params(torch.from_numpy(np.ascontiguousarray(z["params"], dtype=np.float32)).cuda())
    return net, meta


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--samples", type=int, default=1024)
    ap.add_argument("--val", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--n-cell", type=int, default=128)
    ap.add_argument("--level", type=int, default=2)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--body-fraction", type=float, default=0.3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    import torch
    assert torch.cuda.is_available(), "training runs on the GPU"
    nl = a.n_cell >> a.level
    nodes = (nl + 1,) * 3
    _log(f"network level {a.level}: {nodes} nodes; generating {a.samples} + {a.val} samples at "
         f"{a.n_cell + 1}^3 nodes")
    data = generate(a.samples, a.n_cell, a.level, seed0=1000, body_fraction=a.body_fraction)
    val = generate(a.val, a.n_cell, a.level, seed0=500000, body_fraction=a.body_fraction)
    net = UNet(nodes, data[0]["x"].shape[0], 3, seed=a.seed)
    e0 = float(np.median(rel_errors(net, val)))
    _log(f"untrained validation rel L2 {e0:.4f}; {net.num_params} parameters")
    t0 = time.time()
    hist = train(net, data, val, a.epochs, lr_max=a.lr, seed=a.seed)
    torch.cuda.synchronize()
    meta = dict(nodes=list(nodes), n_in=net.n_in, n_out=3, level=a.level, n_cell=a.n_cell, samples=len(data),
                val=len(val), epochs=a.epochs, train_seconds=time.time() - t0, untrained_val_rel_err=e0,
                history=hist, kinds={k: sum(d["kind"].startswith(k) for d in data)
                                     for k in ("uniaxial", "shear", "body")})
    save(a.out, net, meta)
    _log(f"saved {a.out}: final validation rel L2 {hist[-1]['val_rel_err']:.4f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
