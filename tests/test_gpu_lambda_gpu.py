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
