"""Pins of the network oracle (oracle/unet_oracle.py) by mathematics: the
convolution against a direct loop sum, the transfers against their pointwise
definitions and the adjoint identity <up x, y> = <x, 8 down y>, and the whole
backward pass against central finite differences of the float64 forward."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def test_conv3_matches_direct_loops():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((3, 4, 5, 2))
    W = rng.standard_normal((3, 27, 2))
    b = rng.standard_normal(3)
    Y = U.conv3(X, W, b)
    nz, ny, nx = X.shape[:3]
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                ref = b.copy()
                for t, (dz, dy, dx) in enumerate(U.TAPS):
                    zz, yy, xx = z + dz, y + dy, x + dx
                    if 0 <= zz < nz and 0 <= yy < ny and 0 <= xx < nx:
                        ref += W[:, t, :] @ X[zz, yy, xx]
                assert np.allclose(Y[z, y, x], ref, rtol=1e-13, atol=1e-13)


def test_tap_order_is_dz_dy_dx_major():
    assert U.TAPS[0] == (-1, -1, -1) and U.TAPS[13] == (0, 0, 0) and U.TAPS[14] == (0, 0, 1)
    assert all(U.TAPS[26 - t] == tuple(-v for v in U.TAPS[t]) for t in range(27))


def test_up_reproduces_linears_and_down_weights():
    z, y, x = np.meshgrid(np.arange(5.0), np.arange(3.0), np.arange(5.0), indexing="ij")
    lin = (1.0 + 2 * x - 3 * y + 0.5 * z)[..., None]
    fz, fy, fx = np.meshgrid(np.arange(9.0) / 2, np.arange(5.0) / 2, np.arange(9.0) / 2, indexing="ij")
    assert np.allclose(U.up(lin)[..., 0], 1.0 + 2 * fx - 3 * fy + 0.5 * fz)
    # interior full weighting of a constant is the sum of the 27 weights / 8 = 1
    c = np.ones((9, 9, 9, 1))
    d = U.down(c)
    assert np.allclose(d[1:-1, 1:-1, 1:-1], 1.0)
    assert np.isclose(d[0, 0, 0, 0], (1 + 0.5) ** 3 / 8)


def test_up_down_adjoint():
    rng = np.random.default_rng(2)
    xc = rng.standard_normal((3, 5, 4, 2))
    yf = rng.standard_normal((5, 9, 7, 2))
    assert np.isclose(np.sum(U.up(xc) * yf), np.sum(xc * 8.0 * U.down(yf)), rtol=1e-12)


def test_param_count_and_shapes():
    n = U.num_params(27, 3)
    specs = U.conv_specs(27, 3)
    assert len(specs) == 15
    assert specs[8] == (128 + 64, 64, True, 2) and specs[-1] == (16, 3, False, 0)
    assert n == sum(co * 27 * ci + co for ci, co, _, _ in specs)
    assert U.level_shapes((9, 17, 9)) == [(9, 17, 9), (5, 9, 5), (3, 5, 3), (2, 3, 2)]
    with pytest.raises(ValueError):
        U.level_shapes((10, 9, 9))


def test_round_bf16_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0])
    r = U.round_bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5 and r[4] == 0.0


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    n_in, n_out = 2, 3
    p = U.init_params(n_in, n_out, seed=seed) + 0.05 * rng.standard_normal(U.num_params(n_in, n_out))
    x = rng.standard_normal((9, 9, 9, n_in))
    dy = rng.standard_normal((9, 9, 9, n_out))
    g, gx, _ = U.backward(p, x, n_out, dy)

    def L(pp, xx):
        return float(np.sum(U.forward(pp, xx, n_out) * dy))

    # unit directions keep eps * d far below the distance of the ReLU inputs to 0
    eps = 1e-6
    for k in range(3):
        d = rng.standard_normal(p.size)
        d /= np.linalg.norm(d)
        fd = (L(p + eps * d, x) - L(p - eps * d, x)) / (2 * eps)
        assert np.isclose(g @ d, fd, rtol=2e-6, atol=1e-8), (k, g @ d, fd)
    dx = rng.standard_normal(x.shape)
    dx /= np.linalg.norm(dx)
    fd = (L(p, x + eps * dx) - L(p, x - eps * dx)) / (2 * eps)
    assert np.isclose(np.sum(gx * dx), fd, rtol=2e-6, atol=1e-8)
    # single coordinates: the last bias and one weight of the bottleneck
    for idx in (p.size - 1, U.num_params(n_in, n_out) // 2):
        e = np.zeros(p.size)
        e[idx] = 1.0
        fd = (L(p + eps * e, x) - L(p - eps * e, x)) / (2 * eps)
        assert np.isclose(g[idx], fd, rtol=1e-5, atol=1e-8)
