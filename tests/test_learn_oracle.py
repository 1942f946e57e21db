"""CPU checks that pin the learned warm start's oracle (oracle/learn_oracle.py)
and the synthetic scenario helpers."""
import numpy as np
import pytest

import learn_oracle as lo
import lam_synth as syn


def test_bf16_rounding_bit_patterns():
    x = np.array([1.0, -2.5, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 3.0e38, 0.0], np.float32)
    bits = lo.bf16_bits(x)
    # ties round to even: 1 + 2^-8 -> 1.0 (0x3f80), 1 + 3*2^-8 -> 1 + 2^-6 (0x3f82)
    assert list(bits) == [0x3F80, 0xC020, 0x3F80, 0x3F82, 0x7F62, 0x0000]
    assert lo.bf16_value(bits)[1] == -2.5


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    d = lo.dims(n=5, KW=2, H=32)
    M = 6
    params = rng.normal(0, 0.3, d["P"])
    X = np.zeros((M, d["Dp"]))
    X[:, :d["Dx"]] = rng.normal(0, 1, (M, d["Dx"]))
    u = rng.normal(size=(M, 2, 5)) + 1j * rng.normal(size=(M, 2, 5))
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    labels = np.stack([u.real, u.imag], -1)
    K = np.array([0, 1, 2, 2, 1, 2])
    loss, g = lo.loss_and_grad(params, d, X, labels, K, bf16=False)
    # padded weights (input columns >= Dx, output rows >= Dy) carry no gradient
    W1g = g[:d["off_b1"]].reshape(d["H"], d["Dp"])
    assert np.all(W1g[:, d["Dx"]:] == 0)
    idx = rng.choice(np.flatnonzero(np.abs(g) > 1e-6), 40, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        num = (lo.loss_and_grad(pp, d, X, labels, K, bf16=False)[0]
               - lo.loss_and_grad(pm, d, X, labels, K, bf16=False)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_loss_is_phase_invariant_and_zero_at_labels():
    rng = np.random.default_rng(4)
    d = lo.dims(n=4, KW=2, H=32)
    u = rng.normal(size=(3, 2, 4)) + 1j * rng.normal(size=(3, 2, 4))
    u /= np.linalg.norm(u, axis=2, keepdims=True)
    labels = np.stack([u.real, u.imag], -1)
    # a network whose output equals the labels times a phase: W1 = 0, W2 = 0, b2 = labels*phase
    # (one sample at a time, since b2 is shared)
    for m in range(3):
        params = np.zeros(d["P"])
        y = u[m] * np.exp(1j * rng.uniform(0, 2 * np.pi, (2, 1))) * 3.0
        params[d["off_b2"]:d["off_b2"] + d["Dy"]] = np.stack([y.real, y.imag], -1).ravel()
        X = np.zeros((1, d["Dp"]))
        loss, _ = lo.loss_and_grad(params, d, X, labels[m:m + 1], np.array([2]), bf16=False)
        assert loss < 1e-12


def test_features_follow_covariance_and_whitening():
    import oracle as orc
    sh = syn.BlockShape(C=1, F=1, NT=8, R=6, P=2, RPP=6, NP=4, B=1)
    scene = syn.make_scene(sh, seed=1, n_tones=1, amp=30.0)
    ring, scales = syn.make_block(sh, scene, seed=2)
    s = orc.Shape(C=1, F=1, NT=8, R=6, P=2, RPP=6, NP=4, B=1)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)
    A = orc.covariance(6, 2, orc.correlate(s, y)[0], 0)
    Aw, dd = lo.whiten(A)
    Aw_c, dd_c = orc.whiten(A)
    assert np.allclose(Aw, Aw_c, rtol=1e-12) and np.allclose(dd, dd_c, rtol=1e-12)
    d = lo.dims(n=6, KW=2, H=32)
    basis = np.zeros((2, 6), complex)
    basis[0, 1] = 1.0
    bits = lo.features(Aw, basis, d)
    x = lo.bf16_value(bits)
    # entry (i=1, j=3): idx = 3*2/2 + 1 = 4
    assert x[8] == lo.bf16_round(np.float32(Aw[1, 3].real))
    assert x[9] == lo.bf16_round(np.float32(Aw[1, 3].imag))
    off = 6 * 5
    assert x[off + 2] == lo.bf16_round(np.float32(np.sqrt(6)))
    assert np.all(x[d["Dx"]:] == 0)


def test_host_generator_statistics_and_tone():
    sh = syn.BlockShape(C=1, F=2, NT=16, R=8, P=2, RPP=4, NP=8, B=1)
    tones = np.zeros((1, 1, 6), np.float32)
    steer = np.zeros((1, 1, 8, 2), np.float32)
    steer[..., 0] = 1.0
    gain = np.ones(1, np.float32)
    ring, scales = syn.host_synth(sh, 11, 0, 16.0, tones, steer, gain)
    assert abs(ring.astype(np.float64).std() - 16.0) < 0.6
    assert np.all(scales == 1)
    # a pure tone (sigma 0): x = round(A cos(2 pi nu t + phi)) with t = (slot - 1) NT + s
    tones[0, 0] = [50.0, 0.125, 0.3, 1.0, 0.0, 1.0]
    ring, _ = syn.host_synth(sh, 11, 0, 0.0, tones, steer, gain)
    t = 5 * sh.NT + 3  # slot 6, sample 3
    exp0 = np.rint(50 * np.cos(2 * np.pi * 0.125 * t + 0.3))
    assert ring[0, 6, 1, 3, 2, 0, 0] == exp0
    # pol 1 gain i: re(pol1) = -im(pol0 tone)
    assert ring[0, 6, 1, 3, 2, 1, 0] == np.rint(-50 * np.sin(2 * np.pi * 0.125 * t + 0.3))


def test_drift_scenario_switching_and_drift():
    sc = syn.DriftScenario(C=64, R=12, seed=3, n_blocks=50)
    assert sc.nsrc.min() >= 1 and sc.nsrc.max() <= 4
    frac = np.mean(sc.on[1:] != sc.on[:-1], axis=(0,))[sc.exists].mean()
    assert 0.05 < frac < 0.15
    a0, a1 = sc.steering(0), sc.steering(1)
    assert np.allclose(np.abs(a0), 1.0)
    assert not np.allclose(a0, a1)
