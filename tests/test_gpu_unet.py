"""GPU tests of the network kernels (include/alamo_unet.h) against the CPU
oracle oracle/unet_oracle.py with the same bfloat16 rounding points.

Tolerances: the device accumulates in float32 and the oracle in float64, so
a stored bfloat16 value can round to the neighbouring bfloat16 (2^-8
relative) where the two sums straddle a rounding boundary; such flips
propagate.  The bound on the relative L2 difference is 1e-2 for the output
and for each gradient block (observed: output <= 5e-3, gradients <= 3e-4;
the values are printed)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import unet_oracle as U  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def make(nodes, n_in, n_out, seed):
    from alamo_amd.learned import UNet
    rng = np.random.default_rng(seed)
    p = U.init_params(n_in, n_out, seed=seed) + 0.02 * rng.standard_normal(U.num_params(n_in, n_out))
    p = p.astype(np.float32)
    net = UNet(nodes, n_in, n_out, seed=None)
    assert net.num_params == p.size
    net.set_params(torch.from_numpy(p).cuda())
    shape = (nodes[2], nodes[1], nodes[0])
    x = rng.standard_normal(shape + (n_in,)).astype(np.float32)
    return net, p.astype(np.float64), x


def to_dev(a):
    """(nz, ny, nx, C) numpy -> (C, nz, ny, nx) float32 CUDA."""
    return torch.from_numpy(np.ascontiguousarray(np.moveaxis(a, -1, 0)).astype(np.float32)).cuda()


def from_dev(t):
    return np.moveaxis(t.cpu().numpy().astype(np.float64), 0, -1)


# ALAMO_NN_KSPLIT: "1" (the default) splits K over the waves of a workgroup
# on the levels with few nodes, "0" never (every level runs the plain
# kernel's 16-, 32- and 64-channel forms), "2" on every level.
KSPLIT = ["0", "1", "2"]


@pytest.mark.parametrize("ksplit", KSPLIT)
@pytest.mark.parametrize("nodes,n_in", [((9, 9, 9), 3), ((17, 9, 9), 27), ((9, 17, 25), 8)])
def test_forward_matches_oracle(nodes, n_in, ksplit, monkeypatch):
    monkeypatch.setenv("ALAMO_NN_KSPLIT", ksplit)
    net, p, x = make(nodes, n_in, 3, seed=3)
    y = from_dev(net.forward(to_dev(x)))
    torch.cuda.synchronize()
    yo = U.forward(p, U.round_bf16(x), 3, bf16=True)
    e = rel(y, yo)
    print(f"forward {nodes} n_in={n_in} ksplit={ksplit}: rel L2 {e:.2e}")
    assert e < 1e-2


@pytest.mark.parametrize("ksplit", KSPLIT)
@pytest.mark.parametrize("nodes,n_in", [((9, 9, 9), 3), ((17, 9, 9), 27)])
def test_backward_matches_oracle(nodes, n_in, ksplit, monkeypatch):
    """The input gradients of the convolutions run the forward kernels on
    the transposed weights, so the K split reaches them too."""
    monkeypatch.setenv("ALAMO_NN_KSPLIT", ksplit)
    net, p, x = make(nodes, n_in, 3, seed=4)
    rng = np.random.default_rng(5)
    dy = rng.standard_normal((nodes[2], nodes[1], nodes[0], 3))
    net.forward(to_dev(x))
    dp, dx = net.backward(to_dev(dy), want_input_grad=True)
    torch.cuda.synchronize()
    g_o, gx_o, _ = U.backward(p, U.round_bf16(x), 3, U.round_bf16(dy), bf16=True)
    dp = dp.cpu().numpy().astype(np.float64)
    o = 0
    worst = 0.0
    for i, (ci, co, _, _) in enumerate(U.conv_specs(n_in, 3)):
        nw = co * 27 * ci
        ew = rel(dp[o:o + nw], g_o[o:o + nw])
        eb = rel(dp[o + nw:o + nw + co], g_o[o + nw:o + nw + co])
        print(f"conv {i:2d} ({ci}->{co}): dW rel {ew:.2e}, db rel {eb:.2e}")
        worst = max(worst, ew, eb)
        o += nw + co
    ex = rel(from_dev(dx), gx_o)
    print(f"input gradient rel {ex:.2e}")
    assert worst < 1e-2 and ex < 1e-2


def test_mse_and_train_step_decreases_loss():
    nodes, n_in = (9, 9, 9), 4
    net, p, x = make(nodes, n_in, 3, seed=6)
    rng = np.random.default_rng(7)
    t = rng.standard_normal((nodes[2], nodes[1], nodes[0], 3)) * 0.1
    xd, td = to_dev(x), to_dev(t)
    y = net.forward(xd)
    loss, d = net.mse(y, td, want_grad=True)
    yn = y.cpu().numpy()
    tn = td.cpu().numpy()
    assert np.isclose(loss, np.mean((yn - tn) ** 2), rtol=1e-5)
    assert np.allclose(d.cpu().numpy(), 2 * (yn - tn) / yn.size, rtol=1e-5, atol=1e-12)
    losses = [net.train_step(xd, td, 1e-3) for _ in range(30)]
    print("losses", losses[0], losses[-1])
    assert losses[-1] < 0.5 * losses[0]


def test_bad_grid_is_rejected():
    from alamo_amd.elastic import AlamoError
    from alamo_amd.learned import UNet
    with pytest.raises(AlamoError):
        UNet((10, 9, 9), 3)


@pytest.mark.parametrize("trained", [False, True])
def test_backward_on_the_device_tape_matches_oracle_at_33(trained):
    """At the deployed size (33^3 nodes, 51 input channels) bf16 rounding
    flips ReLU masks in the forward pass, and over 15 layers the gradients of
    a free-running oracle drift from the device's by about 8% relative L2.
    Here the oracle's backward runs on the device's own forward tape (each
    convolution's input and ReLU output, alamo_nn_tape), so that the masks
    agree and what is compared is the backward arithmetic: every layer's
    weight and bias gradient and the input gradient within 1e-2 relative L2
    (the values are printed).  With random weights and with the trained ones."""
    import os
    nodes, n_in = (33, 33, 33), 51
    net, p, x = make(nodes, n_in, 3, seed=9)
    if trained:
        from alamo_amd import train
        w = os.path.join(os.path.dirname(train.__file__), "weights", "unet_polycrystal129_level2.npz")
        p = np.load(w)["params"].astype(np.float32)
        net.set_params(torch.from_numpy(p).cuda())
        p = p.astype(np.float64)
    rng = np.random.default_rng(10)
    dy = rng.standard_normal((nodes[2], nodes[1], nodes[0], 3))
    y = from_dev(net.forward(to_dev(x)))
    tape = []
    for i in range(15):
        a, b = net.tape(i)
        tape.append((from_dev(a), from_dev(b) if b is not None else None))
    dp, dx = net.backward(to_dev(dy), want_input_grad=True)
    torch.cuda.synchronize()
    # the tape's first input is the device's bf16 copy of x
    assert np.array_equal(tape[0][0], U.round_bf16(x))
    g_o, gx_o, y_o = U.backward(p, U.round_bf16(x), 3, U.round_bf16(dy), bf16=True, tape_in=tape)
    print(f"forward, free-running oracle: rel L2 {rel(y, y_o):.2e}")
    dp = dp.cpu().numpy().astype(np.float64)
    o = 0
    worst = 0.0
    for i, (ci, co, _, _) in enumerate(U.conv_specs(n_in, 3)):
        nw = co * 27 * ci
        ew = rel(dp[o:o + nw], g_o[o:o + nw])
        eb = rel(dp[o + nw:o + nw + co], g_o[o + nw:o + nw + co])
        print(f"conv {i:2d} ({ci}->{co}): dW rel {ew:.2e}, db rel {eb:.2e}")
        worst = max(worst, ew, eb)
        o += nw + co
    ex = rel(from_dev(dx), gx_o)
    print(f"input gradient rel {ex:.2e}, worst layer {worst:.2e}")
    assert worst < 1e-2 and ex < 1e-2
