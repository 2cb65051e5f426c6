"""HIP policy products against the oracle and float64: the float32 forward is
bit-identical; bf16 MFMA products stay within the bf16 rounding bound
|C - C64| <= 2^-8 sum_k |a_k b_k| (+ float32 slack)."""
import numpy as np
import pytest

import oracle as O
import warporacer3d_amd as R

pytestmark = pytest.mark.gpu


def _policy_pair(dev, beams, hidden, precision=0, seed=42):
    cfg = R.Config(beams=beams, hidden=hidden, precision=precision)
    return R.Policy(dev, seed, cfg), O.Policy(seed, beams, hidden), cfg


# shapes of the fused float32 forward: unrolled (1081/256, 108/64), 16-byte
# staging (cols % 4 == 0) and scalar staging, the generic loop with a partial
# row group (200 beams, H 100), ragged workgroups, and past its column limit
# (20000 columns: the tiled products)
@pytest.mark.parametrize("beams,hidden,cols", [(108, 64, 200), (1081, 256, 300), (13, 7, 5), (1081, 256, 4096),
                                               (108, 64, 1024), (1081, 256, 301), (200, 100, 48), (108, 256, 77),
                                               (1081, 256, 20000)])
def test_init_and_forward_float32_bit_identical(dev, beams, hidden, cols):
    print(f"forward f32 {beams}/{hidden}/{cols}", flush=True)
    g, o, cfg = _policy_pair(dev, beams, hidden)
    assert np.array_equal(g.theta.read(), o.theta), "init differs"
    x = np.random.default_rng(0).uniform(-1, 1, (cfg.IN, cols)).astype(np.float32)
    x[-1] = 1
    xb = dev.upload(x)
    acts = R.Activations(dev, cols, cfg)
    g.forward(R.mat(xb, cfg.IN, cols), acts)
    h1, h2, out = o.forward(x)
    assert np.array_equal(acts.h1.read().reshape(cfg.HA, cols), h1)
    assert np.array_equal(acts.h2.read().reshape(cfg.HA, cols), h2)
    assert np.array_equal(acts.o.read().reshape(3, cols), out)


def test_forward_float32_sees_direct_theta_writes(dev):
    """A write into theta (no transpose) reaches the next forward, as it does
    in the reference, where the forward reads theta itself."""
    beams, hidden, cols = 108, 64, 64
    g, o, cfg = _policy_pair(dev, beams, hidden)
    x = np.random.default_rng(5).uniform(0, 1, (cfg.IN, cols)).astype(np.float32)
    xb = dev.upload(x)
    acts = R.Activations(dev, cols, cfg)
    g.forward(R.mat(xb, cfg.IN, cols), acts)
    o.theta[:] = np.random.default_rng(6).normal(0, 0.1, o.theta.size).astype(np.float32)
    g.theta.write(o.theta)
    g.forward(R.mat(xb, cfg.IN, cols), acts)
    _, _, out = o.forward(x)
    assert np.array_equal(acts.o.read().reshape(3, cols), out)


def test_forward_bf16_within_rounding_bound(dev):
    print("forward bf16 1081/256", flush=True)
    beams, hidden, cols = 1081, 256, 384
    g, o, cfg = _policy_pair(dev, beams, hidden, precision=1)
    x = np.random.default_rng(1).uniform(0, 1, (cfg.IN, cols)).astype(np.float32)
    x[-1] = 1
    xb = dev.upload(x)
    acts = R.Activations(dev, cols, cfg)
    g.forward(R.mat(xb, cfg.IN, cols), acts)
    th = o.theta.astype(np.float64)
    W0 = th[: hidden * cfg.IN].reshape(hidden, cfg.IN)
    pre = W0 @ x.astype(np.float64)
    bound = 2.0**-8 * (np.abs(W0) @ np.abs(x.astype(np.float64))) + 1e-5
    h1 = acts.h1.read().reshape(cfg.HA, cols)
    assert np.all(np.abs(h1[:hidden] - np.tanh(pre)) <= bound)
    assert np.all(h1[hidden] == 1)


def _gemm_case(dev, M, N, K, precision, epi, rng):
    """One call of the policy forward restricted to W0 (M = hidden, K = IN)."""
    beams = K - 5
    cfg = R.Config(beams=beams, hidden=M, precision=precision)
    pol = R.Policy(dev, 3, cfg)
    theta = pol.theta.read()
    W = rng.normal(0, 0.2, (M, K)).astype(np.float32)
    theta[: M * K] = W.ravel()
    pol.theta.write(theta)
    pol.transpose()
    x = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    xb = dev.upload(x)
    acts = R.Activations(dev, N, cfg)
    pol.forward(R.mat(xb, K, N), acts)
    return W.astype(np.float64), x.astype(np.float64), acts.h1.read().reshape(M + 1, N)[:M]


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (17, 100, 37), (256, 513, 1086), (130, 70, 260), (3, 1, 9)])
@pytest.mark.parametrize("precision", [0, 1])
def test_ragged_products(dev, M, N, K, precision):
    print(f"gemm {M}x{N}x{K} p{precision}", flush=True)
    rng = np.random.default_rng(M * 7 + N + K)
    W, x, h = _gemm_case(dev, M, N, K, precision, "tanh", rng)
    ref = np.tanh(W @ x)
    if precision == 0:
        bound = 1e-5 * (np.abs(W) @ np.abs(x)) + 1e-6
    else:
        bound = 2.0**-8 * (np.abs(W) @ np.abs(x)) + 1e-5
    assert np.all(np.abs(h - ref) <= bound)


@pytest.mark.parametrize("precision,cols", [(0, 640), (1, 640), (0, 77), (1, 77)])
def test_backward_against_oracle(dev, precision, cols):
    print(f"backward p{precision} cols={cols}", flush=True)
    beams, hidden = 108, 64
    g, o, cfg = _policy_pair(dev, beams, hidden, precision)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (cfg.IN, cols)).astype(np.float32)
    d3 = rng.normal(0, 0.01, (3, cols)).astype(np.float32)
    xb, d3b = dev.upload(x), dev.upload(d3)
    acts = R.Activations(dev, cols, cfg)
    g.forward(R.mat(xb, cfg.IN, cols), acts)
    d2, d1, grad = dev.alloc(cfg.HA * cols), dev.alloc(cfg.HA * cols), dev.alloc(cfg.params())
    g.backward(R.mat(xb, cfg.IN, cols), acts, R.mat(d3b, 3, cols), d2, d1, grad)
    h1 = acts.h1.read().reshape(cfg.HA, cols)
    h2 = acts.h2.read().reshape(cfg.HA, cols)
    rd2, rd1, rgrad = o.backward(x, h1, h2, d3)
    gd2 = d2.read().reshape(cfg.HA, cols)
    gd1 = d1.read().reshape(cfg.HA, cols)
    gg = grad.read()
    if precision == 0:
        assert np.array_equal(gd2, rd2) and np.array_equal(gd1, rd1)
        assert np.array_equal(gg[: o.LOGSTD], rgrad[: o.LOGSTD])
    else:
        np.testing.assert_array_equal(gd2, rd2)
        scale = np.abs(rgrad[: o.LOGSTD]).max()
        assert np.max(np.abs(gg[: o.LOGSTD] - rgrad[: o.LOGSTD])) <= 2e-2 * scale


def _bf16(a):
    """Round float32 to bf16 (round to nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("beams,hidden,cols", [(1081, 256, 200), (108, 64, 77), (37, 130, 33)])
def test_fused_bf16_forward_and_sample(dev, beams, hidden, cols):
    """The fused bf16 forward (Runner.step's policy half) against a float64
    evaluation of the same bf16-rounded operands; the float32 output layer and
    the sampled actions against the oracle given the kernel's h2 / outputs."""
    print(f"fused bf16 forward {beams}/{hidden}/{cols}", flush=True)
    from helpers import track

    cfg = R.Config(beams=beams, hidden=hidden, precision=1)
    tr = track("ramp10")
    sim = R.Sim(dev, tr, cols, 5, cfg)
    pol = R.Policy(dev, 42, cfg)
    acts = R.Activations(dev, cols, cfg)
    scratch = dev.alloc(2 * cols)
    sim.reset()
    x = sim.obs.read().reshape(cfg.IN, cols).astype(np.float64)
    R._check(R.lib().racer_runner_step(sim.handle, pol.handle, acts.handle, scratch.handle, 0, 9))
    th = pol.theta.read()
    H, IN, HA = hidden, cfg.IN, cfg.HA
    W0 = th[: H * IN].reshape(H, IN)
    W1 = th[H * IN : H * IN + H * HA].reshape(H, HA)
    W2 = th[H * IN + H * HA : H * IN + H * HA + 3 * HA].reshape(3, HA).astype(np.float64)
    h1 = acts.h1.read().reshape(HA, cols)
    h2 = acts.h2.read().reshape(HA, cols)
    o = acts.o.read().reshape(3, cols)
    a0, xb = _bf16(W0), _bf16(x)
    ref1 = np.tanh(a0 @ xb)
    assert np.all(np.abs(h1[:H] - ref1) <= 1e-5 * (np.abs(a0) @ np.abs(xb)) + 2e-6)
    assert np.all(h1[H] == 1) and np.all(h2[H] == 1)
    a1, hb = _bf16(W1), _bf16(h1)
    ref2 = np.tanh(a1 @ hb)
    assert np.all(np.abs(h2[:H] - ref2) <= 1e-5 * (np.abs(a1) @ np.abs(hb)) + 2e-6)
    ref_o = W2 @ h2.astype(np.float64)
    assert np.all(np.abs(o - ref_o) <= 1e-6 * (np.abs(W2) @ np.abs(h2.astype(np.float64))) + 1e-7)
    # the sampled actions follow from o exactly as the oracle's Sim.sample does
    osim = O.Sim(tr, cols, 5, beams=beams)
    oa, ol, ov = np.zeros((2, cols), np.float32), np.zeros(cols, np.float32), np.zeros(cols, np.float32)
    osim.sample(o, th[-2:], oa, ol, ov, False, 9)
    # Runner.step stepped the sim after sampling; its actions buffer holds the sampled actions
    assert np.array_equal(sim.actions.read().reshape(2, cols), oa)
    sc = scratch.read()
    assert np.array_equal(sc[:cols], ol) and np.array_equal(sc[cols:], ov)


@pytest.mark.parametrize("cols", [9000, 640])
def test_bf16_weight_gradients_at_batch_scale(dev, cols):
    """The MFMA weight-gradient products (W0: 256 x 1086, W1: 256 x 257 over K =
    cols, split into 8192-wide chunks) against a float64 product of the same
    bf16-rounded operands, within the float32 accumulation bound
    |G - G64| <= (8192 + chunks) 2^-24 sum_k |a_k b_k|; the float32 output-layer
    gradient (3 rows, the skinny kernel) bit-identical to the oracle."""
    print(f"bf16 weight gradients cols={cols}", flush=True)
    beams, hidden = 1081, 256
    g, o, cfg = _policy_pair(dev, beams, hidden, precision=1)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (cfg.IN, cols)).astype(np.float32)
    x[-1] = 1
    d3 = rng.normal(0, 0.01, (3, cols)).astype(np.float32)
    xb, d3b = dev.upload(x), dev.upload(d3)
    acts = R.Activations(dev, cols, cfg)
    g.forward(R.mat(xb, cfg.IN, cols), acts)
    d2, d1, grad = dev.alloc(cfg.HA * cols), dev.alloc(cfg.HA * cols), dev.alloc(cfg.params())
    g.backward(R.mat(xb, cfg.IN, cols), acts, R.mat(d3b, 3, cols), d2, d1, grad)
    gg = grad.read()
    h1 = acts.h1.read().reshape(cfg.HA, cols)
    h2 = acts.h2.read().reshape(cfg.HA, cols)
    gd1 = d1.read().reshape(cfg.HA, cols)
    gd2 = d2.read().reshape(cfg.HA, cols)
    chunks = -(-cols // 8192)
    slack = (8192 + chunks) * 2.0 ** -24
    for a, b, off, rows, width in ((gd1[:hidden], x, 0, hidden, cfg.IN),
                                   (gd2[:hidden], h1, hidden * cfg.IN, hidden, cfg.HA)):
        a64, b64 = _bf16(a), _bf16(b)
        ref = a64 @ b64.T
        mag = np.abs(a64) @ np.abs(b64).T
        got = gg[off: off + rows * width].reshape(rows, width).astype(np.float64)
        assert np.all(np.abs(got - ref) <= slack * mag + 1e-30), np.max(np.abs(got - ref) / (mag + 1e-30))
    _, _, rgrad = o.backward(x, h1, h2, d3)
    assert np.array_equal(gg[o.W2: o.LOGSTD], rgrad[o.W2: o.LOGSTD])


@pytest.mark.parametrize("big,blocks", [(16384, ((0, 4096), (8192, 4096))),
                                        (65556, ((0, 4096), (61440, 4096), (65536, 20)))])
def test_fused_forward_large_launch_matches_narrow(dev, big, blocks):
    """Large bf16 forwards take other variants of the same arithmetic: 1024
    or more 16-column workgroups the wide one (two-step fragment ring, aliased
    LDS, three workgroups per CU), 65536 or more columns without the sample
    the tiled one (64 columns per workgroup, inputs staged through LDS). Their
    outputs equal, bit for bit, those of narrow launches over the same
    observations; 65556 columns leave the tiled launch a ragged last tile."""
    print(f"fused forward large launch {big}", flush=True)
    cfg = R.Config(beams=1081, hidden=256, precision=1)
    pol = R.Policy(dev, 7, cfg)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1.0, 1.0, (cfg.IN, big)).astype(np.float32)
    x[-1] = 1.0
    xb = dev.alloc(cfg.IN * big)
    xb.write(x.ravel())
    large = R.Activations(dev, big, cfg)
    pol.forward(R.View(xb, cfg.IN, big), large)
    h1, h2, o = (getattr(large, n).read().reshape(-1, big) for n in ("h1", "h2", "o"))
    for c0, n in blocks:
        narrow = R.Activations(dev, n, cfg)
        xs = dev.alloc(cfg.IN * n)
        xs.write(np.ascontiguousarray(x[:, c0:c0 + n]).ravel())
        pol.forward(R.View(xs, cfg.IN, n), narrow)
        for name, ref in (("h1", h1), ("h2", h2), ("o", o)):
            got = getattr(narrow, name).read().reshape(-1, n)
            assert np.array_equal(got, ref[:, c0:c0 + n]), (name, c0)
