"""GPU BF16-MFMA CNN against the FP64 numpy oracle (same BF16-rounded
parameters and inputs).  Tolerances are the BF16 bound: activations are
rounded to 8 significant bits after every layer."""
import numpy as np
import pytest

from distributednavierstokes_amd.cnn import SFSNet
from oracle import oracle_cnn as oc

pytestmark = pytest.mark.gpu


def setup_case(B=2, M=8, seed=0):
    rng = np.random.default_rng(seed)
    net = SFSNet(M=M, batch=B, seed=seed)
    p = (rng.standard_normal(net.nparams) * 0.08).astype(np.float32)
    net.set_params(p)
    x = rng.standard_normal((B, M, M, M, 3)).astype(np.float32)
    t = (rng.standard_normal((B, M, M, M, 6)) * 0.1).astype(np.float32)
    return net, p, x, t


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("M,B", [(8, 2), (16, 2), (32, 1)])
def test_forward_matches_oracle(M, B):
    """M = 32 takes the row-staged LDS convolution, smaller grids the direct one."""
    net, p, x, _ = setup_case(B=B, M=M)
    y = net.forward(x)
    yo, _, _ = oc.forward(oc.bf16_round(p), oc.bf16_round(x))
    assert rel(y, yo) < 2e-2


@pytest.mark.parametrize("M,B", [(8, 2), (16, 2), (32, 1), (64, 1)])
def test_gradients_match_oracle(M, B):
    """M = 8 takes the voxel-tile weight-gradient kernel, M = 16, 32 and 64
    the plane kernel (nine (dy, dx) offsets per block, a ring of three
    staged input rows; x-rows of a multiple of 16); M = 32 and 64 also the
    row-staged forward and data-gradient convolutions, with register
    prefetch at M = 64.  M = 64 is the production grid (the oracle takes
    ~30 s there)."""
    net, p, x, t = setup_case(B=B, M=M, seed=1)
    loss = net.loss_grad(x, t)
    g = net.grads()
    lo, go = oc.loss_grad(oc.bf16_round(p), oc.bf16_round(x), t.astype(np.float64))
    assert abs(loss - lo) < 2e-2 * lo
    off = 0
    for ci, co in oc.CH:
        nw = 27 * co * ci
        assert rel(g[off:off + nw], go[off:off + nw]) < 5e-2
        assert rel(g[off + nw:off + nw + co], go[off + nw:off + nw + co]) < 5e-2
        off += nw + co


@pytest.mark.parametrize("env", [{"DNS_CNN_MFMA": "k8"}, {"DNS_CNN_CONV": "direct"}, {"DNS_CNN_CONV": "rows"},
                                 {"DNS_CNN_WGRAD": "row16"}, {"DNS_CNN_WGRAD": "rows"}])
def test_kernel_variants_match_oracle(env, monkeypatch):
    """The 32x32x8 MFMA kernels (DNS_CNN_MFMA=k8), the direct convolution
    at M = 32 (DNS_CNN_CONV=direct), the row convolution without register
    prefetch (DNS_CNN_CONV=rows; the default takes it below M = 64), and the
    weight gradients by rows with three x-shifted staged copies
    (DNS_CNN_WGRAD=row16) or one staged row per dx triple
    (DNS_CNN_WGRAD=rows): forward and gradients against the oracle."""
    for key, val in env.items():
        monkeypatch.setenv(key, val)
    net, p, x, t = setup_case(B=1, M=32, seed=4)
    y = net.forward(x)
    yo, _, _ = oc.forward(oc.bf16_round(p), oc.bf16_round(x))
    assert rel(y, yo) < 2e-2
    loss = net.loss_grad(x, t)
    g = net.grads()
    lo, go = oc.loss_grad(oc.bf16_round(p), oc.bf16_round(x), t.astype(np.float64))
    assert abs(loss - lo) < 2e-2 * lo
    off = 0
    for ci, co in oc.CH:
        nw = 27 * co * ci
        assert rel(g[off:off + nw], go[off:off + nw]) < 5e-2
        off += nw + co


def test_adam_update():
    net, p, x, t = setup_case(seed=2)
    net.loss_grad(x, t)
    g = net.grads().astype(np.float64)
    net.adam_step(lr=1e-3)
    p1 = net.params()
    m = 0.1 * g
    v = 0.001 * g * g
    want = p - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.abs(p1 - want).max() < 1e-6


def test_training_reduces_loss():
    rng = np.random.default_rng(3)
    net = SFSNet(M=8, batch=2, seed=3)
    x = rng.standard_normal((2, 8, 8, 8, 3)).astype(np.float32)
    t = (rng.standard_normal((2, 8, 8, 8, 6)) * 0.1).astype(np.float32)
    l0 = net.loss_grad(x, t)
    for _ in range(30):
        net.loss_grad(x, t)
        net.adam_step(lr=1e-3)
    assert net.loss_grad(x, t) < 0.7 * l0


def test_lds_weights_match_streamed_weights(monkeypatch):
    """The production convolution at M = 64 with the weights shared through
    LDS (default) and streamed per wavefront (DNS_CNN_WLDS=0): same operands
    in the same order, so forward output, loss and gradients are bitwise
    equal."""
    res = []
    for val in ("1", "0"):
        monkeypatch.setenv("DNS_CNN_WLDS", val)
        net, p, x, t = setup_case(B=1, M=64, seed=5)
        y = net.forward(x)
        loss = net.loss_grad(x, t)
        res.append((y, loss, net.grads()))
        del net
    (y0, l0, g0), (y1, l1, g1) = res
    assert np.array_equal(y0, y1)
    assert l0 == l1
    assert np.array_equal(g0, g1)

