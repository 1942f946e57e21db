"""LambdaGPUPipeline's correlate / beamform / eigen part (LAM_MODE_LAMBDA_GPU)
on the GPU: the reference's EigenvalueBasic golden (tests/test_pipeline.cu:
956-1052) and a scene against the oracle and LAPACK.

Bars: visibilities bit-exact (integer); beams within 1e-3 of the FP64 oracle;
eigenvalues of every pol-pair covariance within 1e-4 of the largest, and the
projector onto the two largest eigenvectors within 1e-4, against
numpy.linalg.eigh of the Hermitian matrix with the oracle's upper triangle."""
import sys

import numpy as np
import pytest

import lam
import lam_synth as syn
import oracle as orc

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bl(i, j):
    return j * (j + 1) // 2 + i


def _o(s):
    return orc.Shape(C=s.C, F=s.F, NT=s.NT, R=s.R, P=s.P, RPP=s.RPP, NP=s.NP, B=s.B)


def rel_err(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def test_eigenvalue_basic_golden():
    """Every sample (2, -2), 4 receivers, 8 samples: each of the four pol-pair
    covariances is 64 everywhere, eigenvalues 0, 0, 0, 256 and principal
    eigenvector 0.5 in every component (the reference asserts both within
    1e-4); the zero eigenspace is degenerate, so its basis is only checked
    for orthonormality."""
    shape = syn.BlockShape(C=1, F=1, NT=8, R=4, P=2, RPP=4, NP=1, B=1, R_pad=32)
    ring, scales = syn.ring_from_fn(shape, lambda *a: (2, -2))
    p = lam.Pipeline(shape, lam.MODE_LAMBDA_GPU)
    p.set_weights(syn.unity_weights(shape)[0])
    out = p.run_host(ring, scales)
    vis = p.dump_visibilities()
    p.close()
    ev = out["evals_out"]
    vec = out["evecs_out"][..., 0] + 1j * out["evecs_out"][..., 1]
    assert np.max(np.abs(ev[..., :3])) <= 1e-4
    assert np.max(np.abs(ev[..., 3] - 256.0)) <= 1e-4
    assert np.max(np.abs(vec[..., 3, :] - 0.5)) <= 1e-4
    for p_ in range(2):
        for q in range(2):
            U = vec[0, p_, q].T  # columns = eigenvectors
            assert np.allclose(U.conj().T @ U, np.eye(4), atol=1e-4)
    assert np.all(out["beams"][0, :, 0] == 8 - 8j)
    assert np.all(vis[0, :, :, :, 0] == 64.0) and np.all(vis[0, :, :, :, 1] == 0.0)


@pytest.mark.parametrize("gains", [False, True])
def test_scene_vs_oracle(gains):
    """Two FPGAs with sub-packet delays (the general loader), a two-tone scene,
    random weights; with gains, A(i, j) = g_ip conj(g_jq) V[bl(i,j)][p][q]."""
    shape = syn.BlockShape(C=3, F=2, NT=64, R=36, P=2, RPP=18, NP=4, B=4)
    _log(f"lambda-gpu scene gains={gains}")
    scene = syn.make_scene(shape, seed=81, n_tones=2, inr_db=(10, 25))
    ring, scales = syn.make_block(shape, scene, seed=82)
    w16, w = syn.random_weights(shape, 83)
    delays = np.array([2, -3], np.int32)
    rng = np.random.default_rng(84)
    g = np.ones((shape.C, 2, shape.R), np.complex64)
    if gains:
        g = (0.5 + rng.uniform(0, 1, g.shape) * np.exp(2j * np.pi * rng.uniform(0, 1, g.shape))).astype(np.complex64)
    p = lam.Pipeline(shape, lam.MODE_LAMBDA_GPU, num_buffers=2)
    p.set_weights(w16)
    p.set_gains(g)
    p.set_subpacket_delays(delays)
    p.set_eig_params(20, 1e-7)
    out = p.run_host(ring, scales)
    vis = p.dump_visibilities_i64()
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE, delays=delays)
    raw = orc.correlate(s, y)
    assert np.array_equal(vis, raw)
    ge = g.astype(np.complex128)
    assert rel_err(out["beams"], orc.beamform(s, y, w * ge[:, :, None, :])) < 1e-3
    evg = out["evals_out"]
    vg = out["evecs_out"][..., 0] + 1j * out["evecs_out"][..., 1]
    n = shape.R
    for c in range(shape.C):
        for p_ in range(2):
            for q in range(2):
                A = np.zeros((n, n), complex)
                for j in range(n):
                    for i in range(j + 1):
                        v = raw[c, bl(i, j), p_, q, 0] + 1j * raw[c, bl(i, j), p_, q, 1]
                        A[i, j] = v * ge[c, p_, i] * np.conj(ge[c, q, j])
                # Hermitian matrix with this upper triangle (real diagonal)
                H = np.triu(A, 1) + np.triu(A, 1).conj().T + np.diag(A.diagonal().real)
                lam_ref, U_ref = np.linalg.eigh(H)
                scale = np.max(np.abs(lam_ref))
                assert np.max(np.abs(evg[c, p_, q] - lam_ref)) <= 1e-4 * scale, (c, p_, q)
                Ug = vg[c, p_, q, -2:].T
                Ur = U_ref[:, -2:]
                assert np.max(np.abs(Ug @ Ug.conj().T - Ur @ Ur.conj().T)) < 1e-4, (c, p_, q)
                # largest-magnitude component of each eigenvector is real and positive
                k = np.argmax(np.abs(vg[c, p_, q]), axis=1)
                top = vg[c, p_, q][np.arange(n), k]
                assert np.all(top.real > 0) and np.max(np.abs(top.imag)) < 1e-6


def test_bandpass_vs_oracle():
    """Bandpass (detect_and_downsample_fft after a block-length FFT): the GPU's
    values of its own fp16 beams within 1e-5 of each row's largest bin, and
    within 2e-3 of the same product of the FP64 oracle's beams (power is the
    square of the beams, so their fp16 rounding counts twice)."""
    shape = syn.BlockShape(C=2, F=1, NT=64, R=8, P=2, RPP=8, NP=8, B=4)
    scene = syn.make_scene(shape, seed=91, n_tones=2)
    ring, scales = syn.make_block(shape, scene, seed=92)
    w16, w = syn.random_weights(shape, 93)
    p = lam.Pipeline(shape, lam.MODE_LAMBDA_GPU)
    p.set_weights(w16)
    p.set_bandpass(16)
    out = p.run_host(ring, scales)
    p.close()
    bp = out["spectra_out"].astype(np.float64)
    own = orc.bandpass(out["beams"].astype(np.complex128), 16)
    assert np.max(np.abs(bp - own) / np.max(own, axis=-1, keepdims=True)) < 1e-5
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)
    fp64 = orc.bandpass(orc.beamform(s, y, w), 16)
    assert rel_err(bp, fp64) < 2e-3


def test_bandpass_8192_points_and_tone():
    """N = 8192 (three radix-16 passes and a radix-2 pass): against numpy's
    FFT of the GPU's beams binned by 128, and a tone at FFT bin 1000 on every
    receiver lands in output bin 1000 // 128 = 7 with (N * 4 * 40)^2 / 128."""
    shape = syn.BlockShape(C=1, F=1, NT=64, R=4, P=2, RPP=4, NP=128, B=2)

    def tone(c, f, pk, t, r, pol):
        a = 2 * np.pi * 1000 * (pk * 64 + t) / 8192
        return (int(round(40 * np.cos(a))), int(round(40 * np.sin(a))))
    ring, scales = syn.ring_from_fn(shape, tone)
    p = lam.Pipeline(shape, lam.MODE_LAMBDA_GPU)
    p.set_weights(syn.unity_weights(shape)[0])
    p.set_bandpass(128)
    out = p.run_host(ring, scales)
    p.close()
    bp = out["spectra_out"].astype(np.float64)
    X = np.fft.fft(out["beams"].astype(np.complex128), axis=-1)
    ref = (np.abs(X) ** 2).reshape(1, 2, 2, 64, 128).mean(-1)
    assert np.max(np.abs(bp - ref) / np.max(ref, axis=-1, keepdims=True)) < 1e-5
    assert np.all(np.argmax(bp, axis=-1) == 7)
    peak = (8192 * 4 * 40.0) ** 2 / 128
    assert np.all(np.abs(bp[..., 7] - peak) < 0.02 * peak)


def test_bandpass_argument_errors():
    shape = syn.BlockShape(C=1, F=1, NT=48, R=4, P=2, RPP=4, NP=1, B=1)  # 48 samples: not a power of two
    p = lam.Pipeline(shape, lam.MODE_LAMBDA_GPU)
    with pytest.raises(lam.LamError):
        p.set_bandpass(16)
    p.close()
    p = lam.Pipeline(syn.BlockShape(C=1, F=1, NT=64, R=4, P=2, RPP=4, NP=1, B=1), lam.MODE_CORR_BEAM)
    with pytest.raises(lam.LamError):
        p.set_bandpass(16)
    p.close()
    # 0 is a no-op outside the LambdaGPU mode: the adaptive spectra keep their buffers
    shape = syn.BlockShape(C=1, F=1, NT=64, R=4, P=2, RPP=4, NP=1, B=1)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
    p.set_weights(syn.unity_weights(shape)[0])
    p.set_fixed_k(np.zeros((1, 2)))
    p.set_spectra(True)
    p.set_bandpass(0)
    ring, scales = syn.ring_from_fn(shape, lambda *a: (2, -2))
    out = p.run_host(ring, scales)
    p.close()
    assert abs(out["spectra_out"][0, 0, 0, 27] - 64 * 4 * np.sqrt(8.0)) < 1e-2


def _herm_upper(A):
    """Hermitian matrix with A's upper triangle (real diagonal), as cuSOLVER's
    fill-UPPER mode reads it."""
    return np.triu(A, 1) + np.triu(A, 1).conj().T + np.diag(A.diagonal().real)


def test_projection_average_vs_numpy():
    """LambdaProjectionPipeline: three blocks, K = 2, averaged every 3 blocks.
    The reference of each (channel, p, q) is built with numpy: per block the
    projector onto the two largest eigenvectors of the covariance (upper
    triangle), averaged, decomposed.  Checked where every block's K-th and
    (K+1)-th eigenvalues are separated by more than 1 % of the largest (else the
    projector itself is ill-defined); the diagonal pol pairs always are."""
    shape = syn.BlockShape(C=2, F=1, NT=64, R=12, P=2, RPP=12, NP=2, B=1)
    scene = syn.make_scene(shape, seed=101, n_tones=2, inr_db=(20, 30))
    blocks = [syn.make_block(shape, scene, seed=110 + b, time_offset=b * shape.T) for b in range(3)]
    p = lam.Pipeline(shape, lam.MODE_PROJECTION, num_buffers=2)
    p.set_projection(2, 3)
    p.set_eig_params(20, 1e-7)
    for ring, scales in blocks:
        p.run_host(ring, scales)
    assert p.projection_runs() == 3
    ev, vec = p.dump_projection()
    assert p.projection_runs() == 0
    with pytest.raises(lam.LamError):
        p.dump_projection()
    p.close()
    s = _o(shape)
    n, K = shape.R, 2
    Pacc = np.zeros((shape.C, 2, 2, n, n), complex)
    sep_ok = np.ones((shape.C, 2, 2), bool)
    for ring, scales in blocks:
        raw = orc.correlate(s, orc.ingest(s, ring, scales, orc.WINDOW_CORR_BEAM))
        for c in range(shape.C):
            for p_ in range(2):
                for q in range(2):
                    A = np.zeros((n, n), complex)
                    for j in range(n):
                        for i in range(j + 1):
                            A[i, j] = raw[c, bl(i, j), p_, q, 0] + 1j * raw[c, bl(i, j), p_, q, 1]
                    lam_b, U = np.linalg.eigh(_herm_upper(A))
                    sep_ok[c, p_, q] &= (lam_b[-K] - lam_b[-K - 1]) > 1e-2 * np.max(np.abs(lam_b))
                    Pacc[c, p_, q] += U[:, -K:] @ U[:, -K:].conj().T
    checked = 0
    for c in range(shape.C):
        for p_ in range(2):
            for q in range(2):
                if p_ == q:
                    assert sep_ok[c, p_, q]
                if not sep_ok[c, p_, q]:
                    continue
                lam_r, U_r = np.linalg.eigh(_herm_upper(Pacc[c, p_, q] / 3))
                assert np.max(np.abs(ev[c, p_, q] - lam_r)) < 1e-4, (c, p_, q)
                Ug = vec[c, p_, q, -K:].T
                assert np.max(np.abs(Ug @ Ug.conj().T - U_r[:, -K:] @ U_r[:, -K:].conj().T)) < 1e-3, (c, p_, q)
                checked += 1
    assert checked >= 4


def test_pulsar_fold_rfi_branch_unwhitened():
    """Pulsar-fold RFI branch (lambda_pulsar_fold_pipeline.hpp:481-640): the
    adaptive path with whitening off.  Eigenvalues of the raw XX / YY
    covariance against LAPACK, and the mitigated beams against W (I - U_K U_K^H)
    applied to the oracle's samples, U_K the two largest LAPACK eigenvectors."""
    shape = syn.BlockShape(C=3, F=1, NT=64, R=16, P=2, RPP=16, NP=4, B=2)
    scene = syn.make_scene(shape, seed=121, n_tones=2, inr_db=(15, 30))
    ring, scales = syn.make_block(shape, scene, seed=122)
    w16, w = syn.random_weights(shape, 123)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
    p.set_weights(w16)
    p.set_fixed_k(np.full((shape.C, 2), 2))
    p.set_whitening(False)
    with pytest.raises(lam.LamError):
        p.set_eig_method(lam.EIG_PREVIOUS)
    p.set_eig_params(20, 1e-7)
    out = p.run_host(ring, scales)
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)
    raw = orc.correlate(s, y)
    n, K = shape.R, 2
    w2 = np.zeros_like(w)
    for c in range(shape.C):
        for p_ in range(2):
            A = np.zeros((n, n), complex)
            for j in range(n):
                for i in range(j + 1):
                    A[i, j] = raw[c, bl(i, j), p_, p_, 0] + 1j * raw[c, bl(i, j), p_, p_, 1]
            H = np.triu(A, 1) + np.triu(A, 1).conj().T + np.diag(A.diagonal().real)
            lam_r, U = np.linalg.eigh(H)
            assert np.max(np.abs(out["evals_out"][c, p_] - lam_r)) <= 1e-4 * lam_r[-1], (c, p_)
            Pm = np.eye(n) - U[:, -K:] @ U[:, -K:].conj().T
            w2[c, p_] = w[c, p_] @ Pm
    assert rel_err(out["beams"][:, :, :2], orc.beamform(s, y, w)) < 1e-3
    assert rel_err(out["beams"][:, :, 2:], orc.beamform(s, y, w2)) < 1e-3


def test_permutation_fused_correlate_beam():
    """LambdaGPU mode with a stream permutation and per-FPGA delays: the
    fused correlate + beam kernel on the general loader (reorder_streams,
    spatial.cuh:593-649); visibilities bit-exact and beams within 1e-3 of the
    oracle, eigenvalues of XX against LAPACK."""
    shape = syn.BlockShape(C=2, F=2, NT=16, R=8, P=2, RPP=4, NP=3, B=2)
    rng = np.random.default_rng(131)
    ring = rng.integers(-50, 51, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    scales = np.ones(shape.scale_shape(), np.int16)
    delays = np.array([3, -5], np.int32)
    recv = np.array([1, 1, 0, 2, -1, 3, 4, 5, 6, 6, 7, 7, 5, 4, 0, 0], np.int32)
    pol = np.array([1, 0, 0, 1, 0, 0, 1, 1, 0, 1, 1, 0, 0, 1, 1, 0], np.int32)
    w16, w = syn.random_weights(shape, 132)
    p = lam.Pipeline(shape, lam.MODE_LAMBDA_GPU)
    p.set_weights(w16)
    p.set_subpacket_delays(delays)
    p.set_stream_permutation(recv, pol)
    p.set_eig_params(20, 1e-7)
    out = p.run_host(ring, scales)
    vis = p.dump_visibilities_i64()
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE, delays=delays, recv_perm=recv, pol_perm=pol)
    raw = orc.correlate(s, y)
    assert np.array_equal(vis, raw)
    assert rel_err(out["beams"], orc.beamform(s, y, w)) < 1e-3
    n = shape.R
    for c in range(shape.C):
        A = np.zeros((n, n), complex)
        for j in range(n):
            for i in range(j + 1):
                A[i, j] = raw[c, bl(i, j), 0, 0, 0] + 1j * raw[c, bl(i, j), 0, 0, 1]
        lam_r = np.linalg.eigvalsh(_herm_upper(A))
        assert np.max(np.abs(out["evals_out"][c, 0, 0] - lam_r)) <= 1e-4 * np.max(np.abs(lam_r))
