"""CPU oracle of the learned coarse-level solver: a 3D U-Net on a nodal grid.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ (and by nothing on the product
path) to check the HIP forward and backward kernels of alamo_amd's network.

The reference has no learned component, so nothing in it pins this module;
it is pinned by mathematics instead (tests/test_unet_oracle.py):
  - conv3 against a direct 7-loop sum,
  - up/down against their matrix forms and the adjoint identity
    <up x, y> = <x, 8 down y>,
  - the whole backward pass against central finite differences of the
    float64 forward pass (rounding off).

Network (channels C = (16, 32, 64, 128), four grid levels, nodal coarsening
n -> (n - 1) / 2 + 1):
  encoder l = 0..3 : e_l = relu(convB(relu(convA(x_l))));  x_0 = input,
                     x_{l+1} = down(e_l)
  decoder l = 2..0 : g_l = relu(convB(relu(convA([up(g_{l+1}), e_l]))));  g_3 = e_3
  output           : y = convO(g_0)          (no activation)
Every conv is 3x3x3, stride 1, zero padding 1, with bias; its weight is stored
W[Cout][27][Cin], tap t = (dz+1)*9 + (dy+1)*3 + (dx+1) reading input node
(x+dx, y+dy, z+dz) (cross-correlation).  down is full weighting scaled by 1/8,
down(X)[I] = 1/8 sum_d w(d) X[2I+d] with w(d) = prod_a (1 - |d_a|/2) and zero
outside the grid; up is trilinear nodal interpolation.

With bf16=True the forward and backward round to bfloat16 (round to nearest
even) at the points the device code stores bfloat16: network input, weights,
every stored activation and every stored activation gradient.  Accumulation
is float64 here (float32 on the device).
"""
from __future__ import annotations

import numpy as np

CHANNELS = (16, 32, 64, 128)
LEVELS = 4


def round_bf16(x):
    """float -> nearest-even bfloat16 -> float64 (NaN-free inputs)."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def level_shapes(nodes):
    """Nodal grid (nz, ny, nx) of each of the LEVELS levels."""
    s = [tuple(int(v) for v in nodes)]
    for _ in range(LEVELS - 1):
        if any((v - 1) % 2 for v in s[-1]) or min(s[-1]) < 3:
            raise ValueError(f"grid {nodes} does not coarsen {LEVELS - 1} times")
        s.append(tuple((v - 1) // 2 + 1 for v in s[-1]))
    return s


def conv_specs(n_in, n_out):
    """(Cin, Cout, relu, level) of the 15 convolutions, in parameter order."""
    c = CHANNELS
    sp = []
    for l in range(LEVELS):
        sp.append((n_in if l == 0 else c[l - 1], c[l], True, l))
        sp.append((c[l], c[l], True, l))
    for l in range(LEVELS - 2, -1, -1):
        sp.append((c[l + 1] + c[l], c[l], True, l))
        sp.append((c[l], c[l], True, l))
    sp.append((c[0], n_out, False, 0))
    return sp


def num_params(n_in, n_out):
    return sum(co * 27 * ci + co for ci, co, _, _ in conv_specs(n_in, n_out))


def split_params(p, n_in, n_out):
    """Flat parameter vector -> [(W[Cout][27][Cin], b[Cout])] in conv order."""
    out, o = [], 0
    for ci, co, _, _ in conv_specs(n_in, n_out):
        W = p[o:o + co * 27 * ci].reshape(co, 27, ci)
        o += co * 27 * ci
        b = p[o:o + co]
        o += co
        out.append((W, b))
    assert o == p.size
    return out


def init_params(n_in, n_out, seed=0):
    """He-normal weights, zero biases (the same rule as the device init)."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co, _, _ in conv_specs(n_in, n_out):
        parts.append(rng.standard_normal(co * 27 * ci) * np.sqrt(2.0 / (27 * ci)))
        parts.append(np.zeros(co))
    return np.concatenate(parts)


TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def _shift(X, d):
    """S[z,y,x] = X[z+dz, y+dy, x+dx], zero outside."""
    nz, ny, nx = X.shape[:3]
    S = np.zeros_like(X)
    dz, dy, dx = d
    zs, ys, xs = slice(max(0, -dz), nz - max(0, dz)), slice(max(0, -dy), ny - max(0, dy)), slice(max(0, -dx), nx - max(0, dx))
    zd, yd, xd = slice(max(0, dz), nz + min(0, dz)), slice(max(0, dy), ny + min(0, dy)), slice(max(0, dx), nx + min(0, dx))
    S[zs, ys, xs] = X[zd, yd, xd]
    return S


def conv3(X, W, b):
    """X (nz, ny, nx, Cin), W (Cout, 27, Cin), b (Cout) -> (nz, ny, nx, Cout)."""
    Y = np.zeros(X.shape[:3] + (W.shape[0],))
    for t, d in enumerate(TAPS):
        Y += _shift(X, d) @ W[:, t, :].T
    return Y + b


def conv3_backward(X, W, dZ):
    """Gradients of conv3 w.r.t. X, W and b for output gradient dZ."""
    dX = np.zeros_like(X)
    dW = np.zeros_like(W)
    Zf = dZ.reshape(-1, dZ.shape[-1])
    for t, d in enumerate(TAPS):
        S = _shift(X, d)
        dW[:, t, :] = Zf.T @ S.reshape(-1, S.shape[-1])
        dX += _shift(dZ @ W[:, t, :], tuple(-v for v in d))
    return dX, dW, Zf.sum(axis=0)


def _w1(d):
    return 1.0 - abs(d) / 2.0


def down(X, scale=0.125):
    """Coarse[I] = scale * sum_d w(d) X[2I+d] (zero outside)."""
    nz, ny, nx = X.shape[:3]
    cz, cy, cx = (nz - 1) // 2 + 1, (ny - 1) // 2 + 1, (nx - 1) // 2 + 1
    Y = np.zeros((cz, cy, cx) + X.shape[3:])
    for d in TAPS:
        w = _w1(d[0]) * _w1(d[1]) * _w1(d[2])
        S = _shift(X, d)
        Y += w * S[::2, ::2, ::2]
    return scale * Y


def up(X):
    """Trilinear nodal interpolation to the grid 2(n-1)+1."""
    cz, cy, cx = X.shape[:3]
    nz, ny, nx = 2 * (cz - 1) + 1, 2 * (cy - 1) + 1, 2 * (cx - 1) + 1
    Y = np.zeros((nz, ny, nx) + X.shape[3:])
    for d in TAPS:
        w = _w1(d[0]) * _w1(d[1]) * _w1(d[2])
        Z = np.zeros_like(Y)
        Z[::2, ::2, ::2] = X
        Y += w * _shift(Z, d)
    return Y


def forward(params, x, n_out, bf16=False, keep=False):
    """x (nz, ny, nx, n_in) -> y (nz, ny, nx, n_out); keep=True also returns
    the tape the backward pass needs."""
    n_in = x.shape[-1]
    r = round_bf16 if bf16 else (lambda v: v)
    layers = [(r(W), b) for W, b in split_params(params, n_in, n_out)]
    specs = conv_specs(n_in, n_out)
    tape = []

    def cv(i, X):
        Y = conv3(X, *layers[i])
        if specs[i][2]:
            Y = np.maximum(Y, 0.0)
            Y = r(Y)
        tape.append((i, X, Y))
        return Y

    h = r(x)
    enc = []
    li = 0
    for l in range(LEVELS):
        h = cv(li, h)
        h = cv(li + 1, h)
        li += 2
        enc.append(h)
        if l < LEVELS - 1:
            h = r(down(h))
    g = enc[-1]
    for l in range(LEVELS - 2, -1, -1):
        u = r(up(g))
        h = np.concatenate([u, enc[l]], axis=-1)
        h = cv(li, h)
        g = cv(li + 1, h)
        li += 2
    y = cv(li, g)
    if keep:
        return y, dict(tape=tape, layers=layers, enc=enc, n_in=n_in)
    return y


def backward(params, x, n_out, dy, bf16=False, tape_in=None):
    """Gradient of <y, dy> w.r.t. the flat parameters and the input x.
    tape_in: optional list, per convolution, of (X, Y) -- its input and its
    output after the ReLU (Y None for the output layer) -- that replaces the
    activations of this forward pass (a device's own forward tape)."""
    n_in = x.shape[-1]
    r = round_bf16 if bf16 else (lambda v: v)
    y, st = forward(params, x, n_out, bf16=bf16, keep=True)
    tape, layers = st["tape"], st["layers"]
    if tape_in is not None:
        tape = [(i, tape_in[i][0], tape_in[i][1] if tape_in[i][1] is not None else Y) for i, _, Y in tape]
    specs = conv_specs(n_in, n_out)
    grads = [None] * len(specs)

    def cv_back(i, dY):
        j, X, Y = tape[i]
        assert j == i
        dZ = dY * (Y > 0) if specs[i][2] else dY
        dZ = r(dZ)
        W = layers[i][0]
        dX, dW, db = conv3_backward(X, W, dZ)
        grads[i] = (dW, db)
        return dX

    li = len(specs) - 1
    dg = r(cv_back(li, dy))
    denc = [None] * LEVELS
    li -= 2
    for l in range(0, LEVELS - 1):
        dh = r(cv_back(li + 1, dg))
        dcat = r(cv_back(li, dh))
        cu = CHANNELS[l + 1]
        du, denc[l] = dcat[..., :cu], dcat[..., cu:]
        dg = r(8.0 * down(du))  # up^T = 8 down
        li -= 2
    denc[LEVELS - 1] = dg
    li = 2 * (LEVELS - 1)
    dh = None
    for l in range(LEVELS - 1, -1, -1):
        de = denc[l]
        if l < LEVELS - 1:
            de = r(de + up(dh) / 8.0)  # down^T = up / 8
        dmid = r(cv_back(li + 1, de))
        dh = r(cv_back(li, dmid))
        li -= 2
    dx = dh
    flat = []
    for dW, db in grads:
        flat.append(dW.ravel())
        flat.append(db)
    return np.concatenate(flat), dx, y
