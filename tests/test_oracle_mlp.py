"""Pin the numpy MLP oracle: its backward pass equals central finite
differences of its own FP64 loss (no BF16 emulation), and the BF16
emulation stays close to FP64."""
import numpy as np
import pytest

from oracle import mlp_ref as M


def random_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.N_PARAMS)
    p[M.OFF_W1:M.OFF_B1] = rng.normal(scale=0.3, size=M.H * M.F)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(scale=0.1, size=M.H)
    p[M.OFF_W2:M.OFF_B2] = rng.normal(scale=0.15, size=M.H * M.H)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(scale=0.1, size=M.H)
    p[M.OFF_W3:M.OFF_B3] = rng.normal(scale=0.2, size=M.H)
    p[M.OFF_B3] = 0.05
    p[M.OFF_MEAN:M.OFF_ISTD] = rng.normal(scale=0.5, size=M.F)
    p[M.OFF_ISTD:M.OFF_OSCALE] = rng.uniform(0.5, 2.0, size=M.F)
    p[M.OFF_OSCALE] = 1.7
    p[M.OFF_ALPHA] = 0.9
    p[M.OFF_PCAP] = 2.5
    p[M.OFF_RESID] = 1.0
    return p


def test_gelu_derivative_matches_finite_difference():
    z = np.linspace(-6, 6, 1001)
    e = 1e-6
    fd = (M.gelu(z + e) - M.gelu(z - e)) / (2 * e)
    assert np.abs(fd - M.dgelu(z)).max() < 1e-8


def test_gelu_limits():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert M.gelu(np.array([10.0]))[0] == pytest.approx(10.0, rel=1e-12)
    assert abs(M.gelu(np.array([-10.0]))[0]) < 1e-12


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = random_params(1)
    x = rng.normal(size=(24, M.F))
    t = rng.normal(size=24)
    _, g = M.loss_and_grad(p, x, t, emulate=False)
    idx = np.concatenate([rng.choice(M.N_TRAIN, 60, replace=False),
                          [M.OFF_B3, M.OFF_W3, M.OFF_B1, M.OFF_W2 + 7]])
    e = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (M.loss_and_grad(pp, x, t, emulate=False)[0] - M.loss_and_grad(pm, x, t, emulate=False)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])
    # statistics (normalisation, scale) are not trained
    assert np.all(g[M.N_TRAIN:] == 0)


@pytest.mark.parametrize("tau", [0.2, 0.8])
def test_expectile_backward_matches_finite_differences(tau):
    rng = np.random.default_rng(2)
    p = random_params(2)
    x = rng.normal(size=(24, M.F))
    t = rng.normal(size=24)
    _, g = M.loss_and_grad(p, x, t, emulate=False, expectile=tau)
    e = 1e-6
    for i in rng.choice(M.N_TRAIN, 40, replace=False).tolist() + [M.OFF_B3]:
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (M.loss_and_grad(pp, x, t, False, tau)[0] - M.loss_and_grad(pm, x, t, False, tau)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_expectile_half_is_mean_squared_error():
    rng = np.random.default_rng(3)
    p = random_params(3)
    x = rng.normal(size=(30, M.F))
    t = rng.normal(size=30)
    la, ga = M.loss_and_grad(p, x, t, emulate=False)
    lb, gb = M.loss_and_grad(p, x, t, emulate=False, expectile=0.5)
    y, _ = M.forward(p, x, emulate=False)
    assert la == lb == pytest.approx(0.5 * np.mean((y - t / p[M.OFF_OSCALE]) ** 2), rel=1e-12)
    assert np.array_equal(ga, gb)


@pytest.mark.parametrize("tau", [0.15, 0.5, 0.9])
def test_expectile_gradient_vanishes_at_the_sample_expectile(tau):
    """A network whose output is the constant b3 (all weights 0) has zero
    b3-gradient exactly at the tau-expectile of the targets: the c with
    tau * sum_{t > c} (t - c) = (1 - tau) * sum_{t < c} (c - t), found by
    bisection.  A swapped weight would put the zero at the 1 - tau expectile."""
    rng = np.random.default_rng(4)
    t = rng.gamma(2.0, size=400)
    lo, hi = t.min(), t.max()
    for _ in range(200):
        c = 0.5 * (lo + hi)
        f = tau * np.sum(np.clip(t - c, 0, None)) - (1 - tau) * np.sum(np.clip(c - t, 0, None))
        lo, hi = (c, hi) if f > 0 else (lo, c)
    p = np.zeros(M.N_PARAMS)
    p[M.OFF_ISTD:M.OFF_OSCALE] = 1.0
    p[M.OFF_OSCALE] = 1.0
    p[M.OFF_B3] = c
    x = rng.normal(size=(t.size, M.F))
    _, g = M.loss_and_grad(p, x, t, emulate=False, expectile=tau)
    assert abs(g[M.OFF_B3]) < 1e-9
    p[M.OFF_B3] = c + 0.01
    _, g = M.loss_and_grad(p, x, t, emulate=False, expectile=tau)
    assert g[M.OFF_B3] > 0


def test_bf16_rounding_is_nearest_even():
    a = np.array([1.0, 1.00390625, 1.005859375, 3.0e38, -2.5], dtype=np.float32)
    r = M.bf16(a)
    assert r[0] == 1.0
    assert r[1] == 1.0            # exactly halfway to the next bf16 -> even (1.0)
    assert r[2] == 1.0078125      # above halfway -> up
    assert r[4] == -2.5


def test_bf16_emulation_close_to_fp64():
    rng = np.random.default_rng(2)
    p = random_params(2)
    x = rng.normal(size=(256, M.F))
    t = rng.normal(size=256)
    y64, _ = M.forward(p, x, emulate=False)
    y16, _ = M.forward(p, x, emulate=True)
    assert np.abs(y64 - y16).max() < 0.05 * np.abs(y64).max()
    _, g64 = M.loss_and_grad(p, x, t, emulate=False)
    _, g16 = M.loss_and_grad(p, x, t, emulate=True)
    assert np.abs(g64 - g16).max() < 0.05 * np.abs(g64).max()


def test_adam_first_step_is_lr_sign():
    """Adam's first bias-corrected step moves each weight by ~lr * sign(g)."""
    p = random_params(3)
    g = np.zeros(M.N_PARAMS)
    g[:M.N_TRAIN] = np.linspace(-1, 1, M.N_TRAIN) + 1e-3
    p2, m, v = M.adam_step(p, g, np.zeros(M.N_TRAIN), np.zeros(M.N_TRAIN), 1, 1e-3)
    step = p[:M.N_TRAIN] - p2[:M.N_TRAIN]
    big = np.abs(g[:M.N_TRAIN]) > 1e-2
    np.testing.assert_allclose(step[big], 1e-3 * np.sign(g[:M.N_TRAIN][big]), rtol=1e-4)


def test_features_beyond_eight_sigma_are_clipped():
    """Normalised inputs are clipped to +-8: a feature 100 standard
    deviations out gives exactly the output of one at 8, and inputs inside
    the range are untouched (the output then depends on them)."""
    p = random_params(3)
    mean, istd = p[M.OFF_MEAN:M.OFF_ISTD], p[M.OFF_ISTD:M.OFF_OSCALE]
    rng = np.random.default_rng(5)
    x = rng.normal(size=(8, M.F))
    far, edge, inside = x.copy(), x.copy(), x.copy()
    far[:, 1] = mean[1] + 100.0 / istd[1]
    edge[:, 1] = mean[1] + 8.0 / istd[1]
    inside[:, 1] = mean[1] + 7.0 / istd[1]
    for emulate in (False, True):
        yf = M.predict(p, far, emulate=emulate)
        ye = M.predict(p, edge, emulate=emulate)
        yi = M.predict(p, inside, emulate=emulate)
        np.testing.assert_array_equal(yf, ye)
        assert np.abs(yi - ye).max() > 0
