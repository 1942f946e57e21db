"""Pins the CPU oracle against the reference's own golden values and against
what the mathematics fixes (LAPACK via scipy, brute force, invariants).

Reference tests cited per case (paths relative to the reference root)."""
import numpy as np
import pytest
import scipy.linalg

import oracle as orc
import lam_synth as syn

SMALL = orc.Shape(C=1, F=1, NT=8, R=4, P=2, RPP=4, NP=1, B=1)       # test_configs.hpp:17-29
TWO_CH = orc.Shape(C=2, F=1, NT=8, R=4, P=2, RPP=4, NP=1, B=1)      # test_configs.hpp:49-61
TWO_PKT = orc.Shape(C=1, F=1, NT=8, R=4, P=2, RPP=4, NP=2, B=1)     # test_configs.hpp:66-78
ADAPT = orc.Shape(C=1, F=1, NT=64, R=4, P=2, RPP=4, NP=1, B=1)      # test_adaptive_pipeline.cu:37-48


def _bs(s):
    return syn.BlockShape(C=s.C, F=s.F, NT=s.NT, R=s.R, P=s.P, RPP=s.RPP, NP=s.NP, B=s.B)


def _corr_beam(shape, sample_fn, scale_fn=None, w=None):
    ring, scales = syn.ring_from_fn(_bs(shape), sample_fn, scale_fn)
    y = orc.ingest(shape, ring, scales, orc.WINDOW_CORR_BEAM)
    if w is None:
        w = np.ones((shape.C, shape.P, shape.B, shape.R), complex)
    return orc.correlate(shape, y), orc.beamform(shape, y, w), y


def bl(i, j):
    """tests/support/assertions.hpp:35-37."""
    return j * (j + 1) // 2 + i


# ---------------------------------------------------------------- correlator
def test_legacy_tcc_golden_conjugation():
    """tests/test_spatial.cpp:77-134: V[bl(4,5)] = sum x4 conj(x5) = 67+29j,
    autos 203 and 190; single sample 23+2j, 13, 41 (test_spatial.cpp:30-75)."""
    x4 = np.array([2 + 3j, 5 + 1j, -2 + 3j, 2 - 3j, 1 + 1j, 0, 10 + 2j, 4 + 4j])
    x5 = np.array([4 + 5j, 2 + 2j, 9 + 1j, -4 - 5j, 0, 1 - 1j, 2 + 2j, 2 + 2j])
    s = orc.Shape(C=1, F=1, NT=8, R=8, P=2, RPP=8, NP=1, B=1)
    y = np.zeros((1, 8, 8, 2, 2), np.int32)
    y[0, :, 4, 0] = np.stack([x4.real, x4.imag], -1)
    y[0, :, 5, 0] = np.stack([x5.real, x5.imag], -1)
    v = orc.correlate(s, y)
    assert tuple(v[0, bl(4, 5), 0, 0]) == (67, 29)
    assert tuple(v[0, bl(4, 4), 0, 0]) == (203, 0)
    assert tuple(v[0, bl(5, 5), 0, 0]) == (190, 0)
    y[:] = 0
    y[0, 0, 4, 0] = (2, 3)
    y[0, 0, 5, 0] = (4, 5)
    v = orc.correlate(s, y)
    assert tuple(v[0, 19, 0, 0]) == (23, 2)
    assert tuple(v[0, 14, 0, 0]) == (13, 0)
    assert tuple(v[0, 20, 0, 0]) == (41, 0)


def test_corr_beam_exact_values():
    """tests/test_corr_beam_pipeline.cu:107-140: beams (8,-8), vis 64+0j."""
    vis, beams, _ = _corr_beam(SMALL, lambda *a: (2, -2))
    assert np.all(beams == 8 - 8j)
    assert np.all(vis[..., 0] == 64) and np.all(vis[..., 1] == 0)


def test_corr_beam_zero_and_zero_weights():
    """test_corr_beam_pipeline.cu:156-211."""
    vis, beams, _ = _corr_beam(SMALL, lambda *a: (0, 0))
    assert not vis.any() and not beams.any()
    vis, beams, _ = _corr_beam(SMALL, lambda *a: (2, -2), w=np.zeros((1, 2, 1, 4)))
    assert not beams.any()
    for r in range(4):
        assert np.all(vis[0, bl(r, r), [0, 1], [0, 1], 0] > 0)


def test_corr_beam_weight_scaling():
    """test_corr_beam_pipeline.cu:213-247: weights 2 -> beams (16,-16)."""
    vis, beams, _ = _corr_beam(SMALL, lambda *a: (2, -2), w=2 * np.ones((1, 2, 1, 4)))
    assert np.all(beams == 16 - 16j)
    assert np.all(vis[..., 0] == 64)


def test_two_channels():
    """test_corr_beam_pipeline.cu:254-299: (8,-8)/(12,4), autos 64/80."""
    vis, beams, _ = _corr_beam(TWO_CH, lambda c, *a: (2, -2) if c == 0 else (3, 1))
    assert np.all(beams[0] == 8 - 8j) and np.all(beams[1] == 12 + 4j)
    for r in range(4):
        for p in range(2):
            assert vis[0, bl(r, r), p, p, 0] == 64 and vis[1, bl(r, r), p, p, 0] == 80


def test_non_unit_scales():
    """test_corr_beam_pipeline.cu:306-337: beam = sum of scales (10, 20),
    autocorrelation N s^2."""
    vis, beams, _ = _corr_beam(SMALL, lambda *a: (1, 0),
                               lambda c, f, k, r, p: (r + 1) * (p + 1))
    assert np.all(beams[0, 0] == 10) and np.all(beams[0, 1] == 20)
    for r in range(4):
        for p in range(2):
            s = (r + 1) * (p + 1)
            assert vis[0, bl(r, r), p, p, 0] == 8 * s * s and vis[0, bl(r, r), p, p, 1] == 0


def test_two_packets():
    """test_corr_beam_pipeline.cu:348-392: 16 samples -> autocorrelation 128."""
    vis, beams, _ = _corr_beam(TWO_PKT, lambda *a: (2, -2))
    assert np.all(vis[0, [bl(r, r) for r in range(4)], 0, 0, 0] == 128)
    assert np.all(beams == 8 - 8j)


def _two_receiver_case(x_rt, w_br):
    """Oracle beams and visibilities of 2 receivers x 2 samples (pol 0)."""
    s = orc.Shape(C=1, F=1, NT=2, R=2, P=2, RPP=2, NP=1, B=2)
    y = np.zeros((1, 2, 2, 2, 2), np.int32)
    x = np.asarray(x_rt)
    y[0, :, :, 0, 0] = x.real.T
    y[0, :, :, 0, 1] = x.imag.T
    w = np.zeros((1, 2, 2, 2), complex)
    w[0, 0] = w_br
    return orc.beamform(s, y, w)[0, 0], orc.correlate(s, y)[0, :, 0, 0]


def test_legacy_beamforming_golden_complex_weights():
    """tests/test_beamforming.cpp:10-81 (legacy, not built): receiver series
    x0 = (1+1j, 2+2j), x1 = (2+2j, 1j), beam weights (1+1j, 1+1j) and
    (2+2j, 2+2j); `weights @ data` -> beams 0+6j, -1+5j / 0+12j, -2+10j with no
    conjugation of the weights; visibilities 10, 6-2j, 9 (x_i conj(x_j))."""
    beams, vis = _two_receiver_case([[1 + 1j, 2 + 2j], [2 + 2j, 1j]], [[1 + 1j, 1 + 1j], [2 + 2j, 2 + 2j]])
    assert np.array_equal(beams, np.array([[6j, -1 + 5j], [12j, -2 + 10j]]))
    assert [tuple(v) for v in vis] == [(10, 0), (6, -2), (9, 0)]


def test_legacy_ccglib_golden_complex_gemm():
    """tests/test_spatial.cpp:221-283 (legacy, commented out): ccglib's
    row-major a times column-major b is a @ b.T; with a = beam weights
    [[1+2j, 2+1j], [-1-2j, 1j]] and b[t][r] = samples [[1+1j, 2+4j], [-2-1j, 2]]
    the beams are [[-1+13j, 4-3j], [-3-1j, 0+7j]]."""
    b = np.array([[1 + 1j, 2 + 4j], [-2 - 1j, 2]])
    beams, _ = _two_receiver_case(b.T, [[1 + 2j, 2 + 1j], [-1 - 2j, 1j]])
    assert np.array_equal(beams, np.array([[-1 + 13j, 4 - 3j], [-3 - 1j, 7j]]))


def test_beamform_brute_force_complex_weights():
    """Random complex weights and int8 samples: oracle == einsum of w x (a
    conjugated or transposed weight operand fails this)."""
    rng = np.random.default_rng(17)
    s = orc.Shape(C=3, F=1, NT=16, R=7, P=2, RPP=7, NP=2, B=5)
    y = rng.integers(-128, 128, (3, 32, 7, 2, 2)).astype(np.int32)
    w = rng.normal(size=(3, 2, 5, 7)) + 1j * rng.normal(size=(3, 2, 5, 7))
    x = y[..., 0].astype(np.float64) + 1j * y[..., 1]
    ref = np.einsum("cpbr,ctrp->cpbt", w, x)
    assert np.allclose(orc.beamform(s, y, w), ref, rtol=1e-12, atol=1e-9)


def test_correlator_brute_force_and_extremes():
    """Random int8 data including -128: oracle == numpy einsum of x x^H."""
    rng = np.random.default_rng(1)
    s = orc.Shape(C=2, F=1, NT=16, R=6, P=2, RPP=6, NP=2, B=2)
    y = rng.integers(-128, 128, (2, 32, 6, 2, 2)).astype(np.int32)
    y[0, 0] = -128
    vis = orc.correlate(s, y)
    x = y[..., 0].astype(np.int64) + 1j * y[..., 1].astype(np.int64)
    full = np.einsum("ctip,ctjq->cijpq", x, np.conj(x))
    for j in range(6):
        for i in range(j + 1):
            ref = full[:, i, j]
            assert np.array_equal(vis[:, bl(i, j), :, :, 0], ref.real.astype(np.int64))
            assert np.array_equal(vis[:, bl(i, j), :, :, 1], ref.imag.astype(np.int64))


def test_window_and_delay_mapping():
    """corr-beam reads ring slots 0..NP-1 (spatial.cuh:254-264); adaptive reads
    slot 1 + per-FPGA sample delay (spatial.cuh:1431-1436)."""
    s = orc.Shape(C=1, F=2, NT=4, R=2, P=2, RPP=1, NP=2, B=1)
    ring = np.zeros((1, 4, 2, 4, 1, 2, 2), np.int8)
    for slot in range(4):
        for f in range(2):
            for t in range(4):
                ring[0, slot, f, t, 0, :, 0] = slot * 10 + t + 50 * f
    scales = np.ones((1, 4, 2, 2), np.int16)
    y = orc.ingest(s, ring, scales, orc.WINDOW_CORR_BEAM)
    assert list(y[0, :, 0, 0, 0]) == [0, 1, 2, 3, 10, 11, 12, 13]
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE, delays=np.array([0, 2]))
    assert list(y[0, :, 0, 0, 0]) == [10, 11, 12, 13, 20, 21, 22, 23]
    assert list(y[0, :, 1, 0, 0]) == [62, 63, 70, 71, 72, 73, 80, 81]
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE, delays=np.array([-4, 0]))
    assert list(y[0, :4, 0, 0, 0]) == [0, 1, 2, 3]


def test_stream_permutation():
    """spatial.cuh:593-649: output (r,p) reads source (recv_perm, pol_perm);
    -1 gives zeros (tests/test_stream_reorder.cu:203-330 pattern)."""
    s = orc.Shape(C=1, F=1, NT=2, R=2, P=2, RPP=2, NP=1, B=1)
    ring = np.zeros((1, 3, 1, 2, 2, 2, 2), np.int8)
    for n in range(2):
        for p in range(2):
            ring[0, :, 0, :, n, p, 0] = 10 * n + p + 1
    scales = np.ones((1, 3, 2, 2), np.int16)
    rp = np.array([1, 1, -1, 0])
    pp = np.array([1, 0, 0, 0])
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE, recv_perm=rp, pol_perm=pp)
    assert list(y[0, 0, :, :, 0].ravel()) == [12, 11, 0, 1]


# ----------------------------------------------------------------- detection
def test_detection_kats():
    """scripts/adaptive_detection_reference.py:86-132 self-test KATs."""
    assert orc.detect_k([0.0, 0.0, 100.0, 100.0], 0.8416) == 0
    assert orc.detect_k(np.zeros(4), 0.5) == 0
    assert orc.detect_k([0.0, 0.0, 0.0, 100.0], 0.5) == 1
    assert orc.detect_k([0.0, 0.0, 100.0, 100.0], 0.5) == 2
    P = orc.null_projection(np.eye(4, dtype=complex), 1)
    assert np.allclose(P, np.diag([1, 1, 1, 0]))
    M = orc.shrink_projection(np.eye(4, dtype=complex), [1.0, 2.0, 3.0, 12.0], 1)
    assert np.allclose(M, np.diag([1, 1, 1, 1 / 6]))


def test_percentile_matches_numpy_linear():
    rng = np.random.default_rng(2)
    v = np.sort(rng.random(37))
    for q in (0.0, 0.2, 0.5, 0.8, 1.0, 0.33):
        assert orc.percentile_linear(v, q) == pytest.approx(np.percentile(v, 100 * q), abs=1e-14)


# ------------------------------------------------------------ whiten / eig
def test_whiten_kats():
    """test_adaptive_pipeline.cu:894-1026."""
    N = 4
    for b in range(2):
        A = np.zeros((N, N), complex)
        for col in range(N):
            for row in range(N):
                A[row, col] = (b + 1) * (row + 1) * (col + 1) + (0.5j if row < col else (-0.5j if row > col else 0))
        W, d = orc.whiten(A)
        assert np.allclose(np.diag(W), 1.0, atol=1e-12)
    A = np.array([[9, 3 - 1j, 6 - 2j], [3 + 1j, 4, 2], [6 + 2j, 2, 16]])
    W, d = orc.whiten(A)
    dd = np.sqrt([9.0, 4.0, 16.0])
    assert np.allclose(W, A / np.outer(dd, dd))
    W, d = orc.whiten(np.zeros((4, 4)))
    assert np.all(np.isfinite(W)) and np.all(np.isfinite(d))


def test_unwhiten_kats():
    """test_adaptive_pipeline.cu:1028-1129."""
    N = 4
    V = np.array([[(1) * (row + 1) - 1j * (col + 1) for col in range(N)] for row in range(N)])
    d = 0.5 * np.arange(1, N + 1)
    U = orc.unwhiten_normalize(V, d)
    assert np.allclose(np.linalg.norm(U, axis=0), 1.0)
    V = np.array([[(col + 1) - 1j * (row + 1) for col in range(N)] for row in range(N)])
    U = orc.unwhiten_normalize(V, np.full(N, 3.0))
    assert np.allclose(U, V / np.linalg.norm(V, axis=0))


def _rand_herm(rng, n, scale=1.0):
    X = rng.standard_normal((n, 3 * n)) + 1j * rng.standard_normal((n, 3 * n))
    return scale * (X @ X.conj().T) / (3 * n)


@pytest.mark.parametrize("n", [1, 2, 4, 12, 24, 36, 64])
def test_jacobi_matches_lapack(n):
    """FP64 Jacobi vs LAPACK zheevd (scipy.linalg.eigh): eigenvalues and
    eigenvector projectors."""
    rng = np.random.default_rng(n)
    A = _rand_herm(rng, n)
    ev, V, sweeps = orc.jacobi_eigh(A)
    ev_ref, V_ref = scipy.linalg.eigh(A)
    assert np.allclose(ev, ev_ref, rtol=1e-12, atol=1e-12)
    assert np.all(np.diff(ev) >= 0)
    assert np.allclose(V.conj().T @ V, np.eye(n), atol=1e-12)
    assert np.allclose(A @ V, V * ev, atol=1e-11)
    # columns equal up to a unit phase
    ph = np.sum(V_ref.conj() * V, axis=0)
    assert np.allclose(np.abs(ph), 1.0, atol=1e-9)


def test_eigenvalue_basic_golden():
    """tests/test_pipeline.cu:956-1052: constant (2,-2) on 4 receivers over 8
    samples -> eigenvalues [0,0,0,256], principal vector entries 0.5."""
    vis, _, _ = _corr_beam(SMALL, lambda *a: (2, -2))
    A = orc.covariance(4, 2, vis[0], 0)
    ev, V, _ = orc.jacobi_eigh(A)
    assert np.allclose(ev, [0, 0, 0, 256], atol=1e-9)
    u = V[:, 3] * np.exp(-1j * np.angle(V[0, 3]))
    assert np.allclose(u, 0.5, atol=1e-12)


def test_projection_algebra():
    rng = np.random.default_rng(5)
    A = _rand_herm(rng, 12)
    ev, V, _ = orc.jacobi_eigh(A)
    P = orc.null_projection(V, 3)
    assert np.allclose(P @ P, P, atol=1e-12)
    assert np.allclose(P @ V[:, -3:], 0, atol=1e-12)
    assert np.allclose(P, np.eye(12) - V[:, -3:] @ V[:, -3:].conj().T)
    M = orc.shrink_projection(V, ev, 0)
    assert np.allclose(M, np.eye(12), atol=1e-12)
    W = rng.standard_normal((4, 12)) + 1j * rng.standard_normal((4, 12))
    assert np.allclose(orc.apply_projection(W, P), W @ P)


def test_fine_delay_table_closed_form():
    """tests/test_fine_delay.cu:56-82 (PhaseTableMath)."""
    tab = orc.fine_delay_table(2, 3, 64, [0.0, 10.0, -5.0], 1.5e9, 10e6, 0)
    for c in range(2):
        for r, dn in enumerate([0.0, 10.0, -5.0]):
            for k in range(64):
                ks = k if k < 32 else k - 64
                phi = -2 * np.pi * (1.5e9 + c * 10e6 + ks / 64 * 10e6) * dn * 1e-9
                assert tab[c, r, k] == pytest.approx(np.exp(1j * phi), abs=1e-12)


# ------------------------------------------------------------ adaptive path
def _adaptive(shape, sample_fn, K=1, shrink=False, detect=False, delta=3.0):
    ring, scales = syn.ring_from_fn(_bs(shape), sample_fn)
    y = orc.ingest(shape, ring, scales, orc.WINDOW_ADAPTIVE)
    w = np.ones((shape.C, shape.P, shape.B, shape.R), complex)
    out = orc.adaptive_block(shape, y, w, fixed_k=np.full((shape.C, shape.P), K),
                             detect=detect, delta=delta, shrink=shrink)
    return out


def test_adaptive_null_rank_one():
    """test_adaptive_pipeline.cu:414-423, 491-511: original (8,-8), mitigated ~0."""
    out = _adaptive(ADAPT, lambda *a: (2, -2), K=1)
    b = out["beams2b"]
    assert np.allclose(b[0, :, 0], 8 - 8j)
    assert np.sum(np.abs(b[0, :, 1]) ** 2) < 1.0
    assert np.all(np.abs(b[0, :, 1]) < 0.25)


def test_adaptive_k0_identity_and_tagged_layout():
    """test_adaptive_pipeline.cu:390-489: K=0 -> beams identical; tagged layout."""
    def tagged(c, f, k, t, r, p):
        return (t + 1, -(1 + 10 * p + 20 * c)) if r == 0 else (0, 0)
    for shrink in (False, True):
        out = _adaptive(ADAPT, tagged, K=0, shrink=shrink)
        b = out["beams2b"]
        for p in range(2):
            exp = np.arange(1, 65) - 1j * (1 + 10 * p)
            assert np.allclose(b[0, p, 0], exp) and np.allclose(b[0, p, 1], exp)


def test_adaptive_shrink_vs_null():
    """test_adaptive_pipeline.cu:622-657: rank-1 shrink == null; rank-2 shrink > null."""
    out = _adaptive(ADAPT, lambda *a: (2, -2), K=1, shrink=True)
    assert np.sum(np.abs(out["beams2b"][0, :, 1]) ** 2) < 1.0

    def rank2(c, f, k, t, r, p):
        if t < 32:
            return (2, -2)
        return (4, 0) if r == 0 else (0, 0)
    pn = np.sum(np.abs(_adaptive(ADAPT, rank2, K=1)["beams2b"][0, :, 1]) ** 2)
    ps = np.sum(np.abs(_adaptive(ADAPT, rank2, K=1, shrink=True)["beams2b"][0, :, 1]) ** 2)
    assert ps > pn + 1e-6


def test_adaptive_detection_counts():
    """test_adaptive_pipeline.cu:659-674: counts [2, 0] with delta 0.5."""
    def multimode(c, f, k, t, r, p):
        if p == 1:
            return (0, 0)
        a = 20 if (t // 8) % 2 == 0 else -20
        b = 18 if (t // 4) % 2 == 0 else -18
        s = 1 if r % 2 == 0 else -1
        return (a + s * b, 0)
    out = _adaptive(ADAPT, multimode, K=0, detect=True, delta=0.5)
    assert list(out["counts"][0]) == [2, 0]
    assert np.sum(np.abs(out["beams2b"][0, 0, 1]) ** 2) < 1.0
    assert np.sum(np.abs(out["beams2b"][0, 1, 1]) ** 2) == 0.0


def test_adaptive_eigen_contract():
    """tests/support/assertions.hpp:136-160: ascending, non-negative."""
    out = _adaptive(ADAPT, lambda *a: (2, -2), K=1)
    ev = out["evals"][0]
    assert np.all(np.diff(ev, axis=-1) >= -1e-9) and np.all(ev >= -1e-9)


def test_null_depth_synthetic_scene():
    """A CW tone from a known direction is nulled by > 20 dB by the K=1
    projection (the physical property the RFI-null path exists for)."""
    shape = syn.BlockShape(C=2, F=1, NT=64, R=12, P=2, RPP=12, NP=4, B=2)
    scene = syn.make_scene(shape, seed=7, n_tones=1, amp=60.0)
    ring, scales = syn.make_block(shape, scene, seed=8)
    s = orc.Shape(C=2, F=1, NT=64, R=12, P=2, RPP=12, NP=4, B=2)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)
    _, w = syn.random_weights(shape, seed=9)
    out = orc.adaptive_block(s, y, w, fixed_k=np.ones((2, 2)))
    for c in range(2):
        a = scene.steering[c, 0]
        for p in range(2):
            # mitigated weight rows are W P; measure their response to a
            W2 = orc.rfi_null_cp(12, 2, out["vis"][c], p, w[c, p], fixed_k=1)[0]
            depth = 10 * np.log10(np.abs(W2[2:] @ a) ** 2 / np.abs(W2[:2] @ a) ** 2)
            assert np.all(depth < -20)


# ------------------------------------------------------------ beam spectra
def _tone(k, n=64, amp=1.0, phase=0.0):
    t = np.arange(n)
    return amp * np.exp(1j * (2 * np.pi * k * t / n + phase))


def test_beam_spectra_tone_closed_form():
    """A complex exponential at DFT bin k has |X[k]| = n*amp and zero
    elsewhere.  Kept bins land at g = k + 27 (k < 27) or k - 37 (k >= 37),
    i.e. DC at output bin 27 and negative frequencies first (adaptive 816-833
    flat copies over the frequency-outermost tensor, 461-464); bins 27..36 are
    dropped.  A sign error in the exponent or a swapped half fails here."""
    for k in range(64):
        x = _tone(k, amp=3.0, phase=0.3 * k)[None, None, None, :]
        out = orc.beam_spectra(x)[0, 0, 0]
        if 27 <= k <= 36:
            assert np.max(out) < 1e-9, k
            continue
        g = k + 27 if k < 27 else k - 37
        assert abs(out[g] - 192.0) < 1e-9, (k, out[g])
        assert np.max(np.delete(out, g)) < 1e-9, k


def test_beam_spectra_packets_sum_and_library_fft():
    """Per-packet magnitudes are summed (sum_fft_over_packets, spatial.cuh:
    1508-1567), and each packet's spectrum equals numpy's FFT."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 2, 3, 4 * 64)) + 1j * rng.standard_normal((2, 2, 3, 4 * 64))
    out = orc.beam_spectra(x)
    X = np.fft.fft(x.reshape(2, 2, 3, 4, 64), axis=-1)
    ref = np.concatenate([np.abs(X[..., 37:]), np.abs(X[..., :27])], axis=-1).sum(axis=3)
    assert np.max(np.abs(out - ref)) < 1e-10 * np.max(ref)
    # packets are independent: spectra of a concatenation are the sum of spectra
    a = orc.beam_spectra(x[..., :128]) + orc.beam_spectra(x[..., 128:])
    assert np.allclose(a, out, rtol=1e-12)
    # |X| is positively homogeneous and phase-blind
    assert np.allclose(orc.beam_spectra(2.5 * np.exp(0.7j) * x), 2.5 * out, rtol=1e-12)


# ------------------------------------------------------ LambdaGPU bandpass
def test_bandpass_tone_and_library_fft():
    """A complex exponential at bin k0 puts (n amp)^2 into that one bin, so
    the run holding it reads (n amp)^2 / d and every other run 0; random rows
    match numpy's FFT binned the same way; a NaN sample leaves no valid value
    in any run (0, as detect_and_downsample_fft writes)."""
    n, d = 256, 16
    t = np.arange(n)
    for k0 in (0, 5, 100, 255):
        x = 1.5 * np.exp(2j * np.pi * k0 * t / n)
        out = orc.bandpass(x[None, :], d)[0]
        exp = np.zeros(n // d)
        exp[k0 // d] = (n * 1.5) ** 2 / d
        assert np.allclose(out, exp, atol=1e-6 * exp.max()), k0
    rng = np.random.default_rng(3)
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    ref = (np.abs(np.fft.fft(x, axis=-1)) ** 2).reshape(3, n // d, d).mean(-1)
    assert np.allclose(orc.bandpass(x, d), ref, rtol=1e-10)
    y = x[:1].copy()
    y[0, 7] = np.nan
    assert np.all(orc.bandpass(y, d) == 0.0)
