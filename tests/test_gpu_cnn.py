"""The closure network's HIP kernels (csrc/closure.hip: MFMA bf16 convolutions, weight gradients,
the last layer, MSE, Adam) against the CPU oracle (oracle/cnn_oracle.py). With the oracle rounding
to bf16 where the kernels store bf16, the remaining difference is fp32 against float64
accumulation, which can move an occasional value by one bf16 step: outputs and gradients are
compared normwise at 5e-3 / 2e-2. Against the pure float64 network the difference is the bf16
error itself (a few 1e-3)."""
import os
import sys

import numpy as np
import pytest

import nse2d_hip as nh

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cnn_oracle as co  # noqa: E402

pytestmark = pytest.mark.gpu


def nrel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def make(H, W, B, seed=5):
    c = nh.lib().nse2d_cnn_create(H, W, B, seed, None)
    assert c, nh.last_error()
    n = nh.lib().nse2d_cnn_num_params(c)
    p = np.zeros(n, np.float32)
    nh.lib().nse2d_cnn_get_params(c, p.ctypes.data)
    return c, p


# 33x512x4 has 1056 row segments of 64 pixels: more than the last layer's gradient kernels take
# in one pass (1024 workgroups), so their segment loops run more than once
@pytest.mark.parametrize("H,W,B", [(9, 64, 2), (7, 128, 1), (33, 512, 4)])
def test_forward_and_gradients_match_oracle(H, W, B):
    print(f"[cnn] {B}x{H}x{W}", flush=True)
    c, p = make(H, W, B)
    lay, n = co.layout()
    assert p.size == n
    rng = np.random.default_rng(H * W)
    inp = rng.normal(size=(B, H, W, co.NIN)).astype(np.float32)
    d_in, d_out = nh.DeviceArray(inp), nh.DeviceArray(shape=(B, H, W))
    assert nh.lib().nse2d_cnn_forward(c, d_in.ptr, d_out.ptr, B) == 0
    out = d_out.download().astype(np.float64)
    ref, cache = co.forward(p.astype(np.float64), inp.astype(np.float64), emulate=True)
    ref64, cache64 = co.forward(p.astype(np.float64), inp.astype(np.float64), emulate=False)
    e, e64 = nrel(out, ref), nrel(out, ref64)
    print(f"[cnn] forward: vs bf16 oracle {e:.2e}, vs float64 {e64:.2e}", flush=True)
    assert e <= 5e-3 and e64 <= 5e-2
    gout = rng.normal(size=(B, H, W)).astype(np.float32)
    d_g = nh.DeviceArray(gout)
    assert nh.lib().nse2d_cnn_backward(c, d_g.ptr, B) == 0
    g = np.zeros(n, np.float32)
    nh.lib().nse2d_cnn_get_grads(c, g.ctypes.data)
    gr = co.backward(p.astype(np.float64), cache, gout.astype(np.float64), emulate=True)
    gr64 = co.backward(p.astype(np.float64), cache64, gout.astype(np.float64), emulate=False)
    for l, (ow, cout, cin, ob) in enumerate(lay):
        sl = slice(ow, ow + 9 * cout * cin)
        if l == 0:  # the padded input channels carry zero gradient
            gw = g[sl].reshape(9, cout, cin)
            assert np.all(gw[..., co.NIN:] == 0)
        ew, eb = nrel(g[sl], gr[sl]), nrel(g[ob:ob + cout], gr[ob:ob + cout])
        ew64 = nrel(g[sl], gr64[sl])
        print(f"[cnn] layer {l}: dW {ew:.2e} (float64 {ew64:.2e}) db {eb:.2e}", flush=True)
        assert ew <= 2e-2 and eb <= 2e-2 and ew64 <= 1e-1
    nh.lib().nse2d_cnn_destroy(c)


def test_adam_step_matches_oracle():
    c, p = make(5, 64, 1)
    rng = np.random.default_rng(9)
    inp = rng.normal(size=(1, 5, 64, co.NIN)).astype(np.float32)
    d_in = nh.DeviceArray(inp)
    nh.lib().nse2d_cnn_forward(c, d_in.ptr, None, 1)
    d_g = nh.DeviceArray(rng.normal(size=(1, 5, 64)).astype(np.float32))
    nh.lib().nse2d_cnn_backward(c, d_g.ptr, 1)
    g = np.zeros(p.size, np.float32)
    nh.lib().nse2d_cnn_get_grads(c, g.ctypes.data)
    m = v = np.zeros(p.size)
    q = p.astype(np.float64)
    for step in (1, 2):
        assert nh.lib().nse2d_cnn_adam_step(c, 1e-3) == 0
        q, m, v = co.adam(q, m, v, g.astype(np.float64), step, 1e-3)
    got = np.zeros(p.size, np.float32)
    nh.lib().nse2d_cnn_get_params(c, got.ctypes.data)
    assert np.max(np.abs(got - q)) <= 1e-6 * max(1.0, np.max(np.abs(q)))
    nh.lib().nse2d_cnn_destroy(c)


def test_training_reduces_loss_on_a_fixed_batch():
    H, W, B = 17, 64, 4
    c, p = make(H, W, B, seed=11)
    rng = np.random.default_rng(12)
    inp = rng.normal(size=(B, H, W, co.NIN)).astype(np.float32)
    tgt = (0.5 * inp[..., 0] - 0.3 * np.roll(inp[..., 1], 1, axis=2)).astype(np.float32)
    d_in, d_t = nh.DeviceArray(inp), nh.DeviceArray(tgt)
    losses = [nh.lib().nse2d_cnn_train_step(c, d_in.ptr, d_t.ptr, B, 1e-3) for _ in range(60)]
    print(f"[cnn] loss {losses[0]:.3e} -> {losses[-1]:.3e}", flush=True)
    out, _ = co.forward(p.astype(np.float64), inp.astype(np.float64), emulate=True)
    assert abs(losses[0] - co.mse(out, tgt)[0]) <= 1e-2 * losses[0]
    assert losses[-1] < 0.2 * losses[0]
    nh.lib().nse2d_cnn_destroy(c)


def test_last_layer_gradient_identities_at_full_size():
    """The last layer is linear in its weights and bias, so for any output gradient g:
    sum_p g[p] (out[p] - b) = <dW, w> and db = sum_p g[p]. With g = out - b the left side is a sum
    of squares (no cancellation), and a weight gradient that gathers any tap, edge column or edge
    row from other pixels than the forward does leaves the identity by far more than the fp32
    rounding of the two sums. At 257x512x2 (4112 row segments, several per workgroup)."""
    H, W, B = 257, 512, 2
    print(f"[cnn] last-layer identities {B}x{H}x{W}", flush=True)
    c, p = make(H, W, B, seed=3)
    lay, n = co.layout()
    ow, cout, cin, ob = lay[-1]
    rng = np.random.default_rng(17)
    inp = rng.normal(size=(B, H, W, co.NIN)).astype(np.float32)
    d_in, d_out = nh.DeviceArray(inp), nh.DeviceArray(shape=(B, H, W))
    assert nh.lib().nse2d_cnn_forward(c, d_in.ptr, d_out.ptr, B) == 0
    out = d_out.download()
    gout = (out - p[ob]).astype(np.float32)
    d_g = nh.DeviceArray(gout)
    assert nh.lib().nse2d_cnn_backward(c, d_g.ptr, B) == 0
    g = np.zeros(n, np.float32)
    nh.lib().nse2d_cnn_get_grads(c, g.ctypes.data)
    nh.lib().nse2d_cnn_destroy(c)
    w = p[ow:ow + 9 * cout * cin].astype(np.float64)
    dw = g[ow:ow + 9 * cout * cin].astype(np.float64)
    g64 = gout.astype(np.float64)
    lhs = float(np.sum(g64 * (out.astype(np.float64) - float(p[ob]))))
    rhs = float(np.dot(dw, w))
    print(f"[cnn] <g, out - b> {lhs:.8e}, <dW, w> {rhs:.8e}, rel {abs(lhs - rhs) / lhs:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-6 * lhs
    assert abs(float(g[ob]) - float(np.sum(g64))) <= 1e-5 * float(np.sum(np.abs(g64)))


@pytest.mark.parametrize("H,W,B", [(9, 64, 2), (7, 128, 1), (33, 64, 3), (257, 512, 2)])
def test_staged_convolution_matches_direct(H, W, B):
    """The LDS-staged convolution accumulates the same products in the same order as the kernel
    that reads its operands from global memory (taps beyond the walls add exact zeros), so the
    network outputs and gradients agree bit for bit; inference (no pre-activations kept) gives
    the same outputs."""
    print(f"[cnn] staged vs direct {B}x{H}x{W}", flush=True)
    c, p = make(H, W, B, seed=9)
    rng = np.random.default_rng(H + W)
    inp = rng.normal(size=(B, H, W, co.NIN)).astype(np.float32)
    gout = rng.normal(size=(B, H, W)).astype(np.float32)
    d_in, d_out, d_g = nh.DeviceArray(inp), nh.DeviceArray(shape=(B, H, W)), nh.DeviceArray(gout)
    res = {}
    for direct in (1, 0, 2):  # global operands, persistent LDS-staged (default), per-tile LDS-staged
        assert nh.lib().nse2d_cnn_set_direct(c, direct) == 0
        assert nh.lib().nse2d_cnn_forward(c, d_in.ptr, d_out.ptr, B) == 0
        out = d_out.download()
        assert nh.lib().nse2d_cnn_backward(c, d_g.ptr, B) == 0
        g = np.zeros(p.size, np.float32)
        nh.lib().nse2d_cnn_get_grads(c, g.ctypes.data)
        assert nh.lib().nse2d_cnn_infer(c, d_in.ptr, d_out.ptr, B) == 0
        res[direct] = (out, g, d_out.download())
    nh.lib().nse2d_cnn_destroy(c)
    for m in (0, 2):
        assert np.array_equal(res[m][0], res[1][0]), m
        assert np.array_equal(res[m][1], res[1][1]), m
    for m in (0, 1, 2):
        assert np.array_equal(res[m][2], res[m][0]), m


def test_forward_and_gradients_at_the_les_workload_on_a_subsample(progress):
    """The network at the learned-closure workload (257x512 images, batch 32): the GPU runs the
    forward and the backward of the whole batch; the output gradient is nonzero only on three row
    bands (the two walls and the middle) of two images, so the oracle needs only those images,
    cropped to each band with 16 rows of margin (the network sees 6 rows up and down; activations
    within 10 rows of a cut edge are exact): outputs on the bands and every parameter gradient
    (the sum over the crops) against the bf16-emulating oracle, at the bars of the small-size
    tests."""
    H, W, B = 257, 512, 32
    progress(f"[cnn] workload {B}x{H}x{W}, oracle on a subsample")
    c, p = make(H, W, B, seed=21)
    lay, n = co.layout()
    rng = np.random.default_rng(257)
    inp = rng.normal(size=(B, H, W, co.NIN)).astype(np.float32)
    bands = [(0, 6), (124, 132), (251, 257)]
    images = (0, 17)
    gout = np.zeros((B, H, W), np.float32)
    for i in images:
        for b0, b1 in bands:
            gout[i, b0:b1] = rng.normal(size=(b1 - b0, W))
    d_in, d_out, d_g = nh.DeviceArray(inp), nh.DeviceArray(shape=(B, H, W)), nh.DeviceArray(gout)
    assert nh.lib().nse2d_cnn_forward(c, d_in.ptr, d_out.ptr, B) == 0
    out = d_out.download().astype(np.float64)
    assert nh.lib().nse2d_cnn_backward(c, d_g.ptr, B) == 0
    g = np.zeros(n, np.float32)
    nh.lib().nse2d_cnn_get_grads(c, g.ctypes.data)
    nh.lib().nse2d_cnn_destroy(c)
    p64 = p.astype(np.float64)
    gr = np.zeros(n)
    got_o, ref_o = [], []
    for i in images:
        for b0, b1 in bands:
            lo, hi = max(0, b0 - 16), min(H, b1 + 16)
            ref, cache = co.forward(p64, inp[i:i + 1, lo:hi].astype(np.float64), emulate=True)
            got_o.append(out[i, b0:b1])
            ref_o.append(ref[0, b0 - lo:b1 - lo])
            gr += co.backward(p64, cache, gout[i:i + 1, lo:hi].astype(np.float64), emulate=True)
    e = nrel(np.concatenate(got_o), np.concatenate(ref_o))
    progress(f"[cnn] workload forward on the bands vs bf16 oracle {e:.2e}")
    assert e <= 5e-3
    for l, (ow, cout, cin, ob) in enumerate(lay):
        sl = slice(ow, ow + 9 * cout * cin)
        ew, eb = nrel(g[sl], gr[sl]), nrel(g[ob:ob + cout], gr[ob:ob + cout])
        progress(f"[cnn] workload layer {l}: dW {ew:.2e} db {eb:.2e}")
        assert ew <= 2e-2 and eb <= 2e-2
