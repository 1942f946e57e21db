"""GPU parity of the HIP path against the CPU oracle and the reference's own
golden values, through the C ABI (liblam.so).

Bars: visibilities bit-exact (integer); beams within 1e-3 normwise relative
error of the FP64 oracle (fp16 output rounding included); eigenvalues of the
whitened covariance within 1e-4 relative to the largest; projections within
1e-4.  Reference tests cited per case."""
import sys

import numpy as np
import pytest

import lam
import lam_synth as syn
import oracle as orc

pytestmark = pytest.mark.gpu

BEAM_TOL = 1e-3


def _o(s):
    return orc.Shape(C=s.C, F=s.F, NT=s.NT, R=s.R, P=s.P, RPP=s.RPP, NP=s.NP, B=s.B)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bl(i, j):
    return j * (j + 1) // 2 + i


def rel_err(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


SMALL = syn.BlockShape(C=1, F=1, NT=8, R=4, P=2, RPP=4, NP=1, B=1, R_pad=32)
TWO_CH = syn.BlockShape(C=2, F=1, NT=8, R=4, P=2, RPP=4, NP=1, B=1, R_pad=32)
TWO_PKT = syn.BlockShape(C=1, F=1, NT=8, R=4, P=2, RPP=4, NP=2, B=1, R_pad=32)
ADAPT = syn.BlockShape(C=1, F=1, NT=64, R=4, P=2, RPP=4, NP=1, B=1, R_pad=32)


def corr_beam_run(shape, sample_fn, scale_fn=None, weights=None, nblocks=1):
    _log(f"corr-beam run {shape}")
    ring, scales = syn.ring_from_fn(shape, sample_fn, scale_fn)
    p = lam.Pipeline(shape, lam.MODE_CORR_BEAM)
    w16 = syn.unity_weights(shape)[0] if weights is None else weights
    p.set_weights(w16)
    for _ in range(nblocks):
        out = p.run_host(ring, scales)
    vis = p.dump_visibilities()
    p.close()
    return out["beams"], vis, ring, scales


# ------------------------------------------------ reference goldens (corr-beam)
def test_corr_beam_exact_values():
    """tests/test_corr_beam_pipeline.cu:107-140."""
    beams, vis, _, _ = corr_beam_run(SMALL, lambda *a: (2, -2))
    assert np.all(beams == 8 - 8j)
    assert np.all(vis[..., 0] == 64.0) and np.all(vis[..., 1] == 0.0)


def test_corr_beam_zero_and_weights():
    """test_corr_beam_pipeline.cu:156-247."""
    beams, vis, _, _ = corr_beam_run(SMALL, lambda *a: (0, 0))
    assert not beams.any() and not vis.any()
    zero_w = np.zeros((1, 2, 1, 4, 2), np.float16).view(np.uint16)
    beams, vis, _, _ = corr_beam_run(SMALL, lambda *a: (2, -2), weights=zero_w)
    assert not beams.any()
    assert np.all(vis[0, [bl(r, r) for r in range(4)], 0, 0, 0] > 0)
    two_w = syn.unity_weights(SMALL, 2.0)[0]
    beams, vis, _, _ = corr_beam_run(SMALL, lambda *a: (2, -2), weights=two_w)
    assert np.all(beams == 16 - 16j) and np.all(vis[..., 0] == 64)


def test_corr_beam_two_channels_scales_packets():
    """test_corr_beam_pipeline.cu:254-392."""
    beams, vis, _, _ = corr_beam_run(TWO_CH, lambda c, *a: (2, -2) if c == 0 else (3, 1))
    assert np.all(beams[0] == 8 - 8j) and np.all(beams[1] == 12 + 4j)
    for r in range(4):
        assert vis[0, bl(r, r), 0, 0, 0] == 64 and vis[1, bl(r, r), 1, 1, 0] == 80
    beams, vis, _, _ = corr_beam_run(SMALL, lambda *a: (1, 0), lambda c, f, k, r, p: (r + 1) * (p + 1))
    assert np.all(beams[0, 0] == 10) and np.all(beams[0, 1] == 20)
    for r in range(4):
        for p in range(2):
            s = (r + 1) * (p + 1)
            assert vis[0, bl(r, r), p, p, 0] == 8 * s * s
    beams, vis, _, _ = corr_beam_run(TWO_PKT, lambda *a: (2, -2))
    assert np.all(vis[0, [bl(r, r) for r in range(4)], 0, 0, 0] == 128)


def test_legacy_tcc_golden_conjugation():
    """tests/test_spatial.cpp:77-134 through the GPU correlator: 67+29j."""
    x4 = [2 + 3j, 5 + 1j, -2 + 3j, 2 - 3j, 1 + 1j, 0, 10 + 2j, 4 + 4j]
    x5 = [4 + 5j, 2 + 2j, 9 + 1j, -4 - 5j, 0, 1 - 1j, 2 + 2j, 2 + 2j]
    shape = syn.BlockShape(C=1, F=1, NT=8, R=8, P=2, RPP=8, NP=1, B=1)

    def fn(c, f, k, t, r, p):
        if p != 0 or k != -1:  # corr-beam reads ring slot 0 = packet -1
            return (0, 0)
        v = x4[t] if r == 4 else (x5[t] if r == 5 else 0)
        return (int(np.real(v)), int(np.imag(v)))
    _, vis, _, _ = corr_beam_run(shape, fn)
    assert tuple(vis[0, bl(4, 5), 0, 0]) == (67.0, 29.0)
    assert tuple(vis[0, bl(4, 4), 0, 0]) == (203.0, 0.0)
    assert tuple(vis[0, bl(5, 5), 0, 0]) == (190.0, 0.0)


def _golden_weights(shape, w_br):
    """FP16 weights [C][P][B][R][2] with beam rows w_br on receivers 0..1 of every (c, p)."""
    halfs = np.zeros((shape.C, shape.P, shape.B, shape.R, 2), np.float16)
    w = np.asarray(w_br)
    halfs[..., :2, 0] = w.real
    halfs[..., :2, 1] = w.imag
    return halfs.view(np.uint16)


@pytest.mark.parametrize("case", ["beamforming", "ccglib"])
@pytest.mark.parametrize("mode", [lam.MODE_CORR_BEAM, lam.MODE_ADAPTIVE])
def test_complex_weight_goldens(case, mode):
    """Complex weights, no conjugation: tests/test_beamforming.cpp:55-80 (beams
    0+6j, -1+5j / 0+12j, -2+10j; visibilities 10, 6-2j, 9) and the ccglib
    a @ b.T golden of tests/test_spatial.cpp:221-283 (-1+13j, 4-3j / -3-1j, 7j),
    through the corr-beam kernel and the adaptive 2B-beam kernel (K = 0, so the
    mitigated beams equal the originals)."""
    shape = syn.BlockShape(C=1, F=1, NT=8, R=4, P=2, RPP=4, NP=1, B=2, R_pad=32)
    if case == "beamforming":
        x = [[1 + 1j, 2 + 2j], [2 + 2j, 1j]]            # x[receiver][time]
        w = [[1 + 1j, 1 + 1j], [2 + 2j, 2 + 2j]]        # w[beam][receiver]
        want = np.array([[6j, -1 + 5j], [12j, -2 + 10j]])
    else:
        x = np.array([[1 + 1j, 2 + 4j], [-2 - 1j, 2]]).T
        w = [[1 + 2j, 2 + 1j], [-1 - 2j, 1j]]
        want = np.array([[-1 + 13j, 4 - 3j], [-3 - 1j, 7j]])
    x = np.asarray(x)

    def fn(c, f, k, t, r, p):
        return (int(x[r, t].real), int(x[r, t].imag)) if (r < 2 and t < 2) else (0, 0)

    _log(f"complex-weight golden {case} mode {mode}")
    ring, scales = syn.ring_from_fn(shape, fn)
    p = lam.Pipeline(shape, mode)
    p.set_weights(_golden_weights(shape, w))
    if mode == lam.MODE_ADAPTIVE:
        p.set_fixed_k(np.zeros((1, 2), np.int32))
    out = p.run_host(ring, scales)
    vis = p.dump_visibilities()
    p.close()
    b = out["beams"]
    for pol in range(2):
        for rep in range(b.shape[2] // 2):
            assert np.array_equal(b[0, pol, 2 * rep:2 * rep + 2, :2], want)
            assert not b[0, pol, 2 * rep:2 * rep + 2, 2:].any()
    if case == "beamforming":
        for pq in [(0, 0), (0, 1), (1, 0), (1, 1)]:
            assert [tuple(v) for v in vis[0, :3, pq[0], pq[1]]] == [(10, 0), (6, -2), (9, 0)]


# ------------------------------------------------------- random-data parity
@pytest.mark.parametrize("shape", [
    syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=4, B=4),
    syn.BlockShape(C=3, F=4, NT=16, R=36, P=2, RPP=9, NP=5, B=2),
    syn.BlockShape(C=2, F=3, NT=8, R=6, P=2, RPP=2, NP=3, B=1),
    syn.BlockShape(C=2, F=2, NT=32, R=24, P=2, RPP=12, NP=4, B=16),
    syn.BlockShape(C=1, F=1, NT=64, R=64, P=2, RPP=64, NP=2, B=8),
], ids=lambda s: f"C{s.C}F{s.F}NT{s.NT}R{s.R}NP{s.NP}B{s.B}")
def test_corr_beam_random_bit_exact(shape):
    """Random int8 (incl. -128): visibilities bit-exact over 2 accumulated
    blocks, beams within 1e-3 of the FP64 oracle."""
    _log(f"random corr-beam {shape}")
    rng = np.random.default_rng(shape.C * 1000 + shape.R)
    ring = rng.integers(-128, 128, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    ring.reshape(-1)[::97] = -128
    scales = np.ones(shape.scale_shape(), np.int16)
    w16, w = syn.random_weights(shape, seed=3)
    p = lam.Pipeline(shape, lam.MODE_CORR_BEAM, num_buffers=2)
    p.set_weights(w16)
    out1 = p.run_host(ring, scales)
    out2 = p.run_host(ring, scales)
    vis64 = p.dump_visibilities_i64()
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_CORR_BEAM)
    ref_vis = orc.correlate(s, y)
    assert np.array_equal(vis64, 2 * ref_vis)
    ref_beams = orc.beamform(s, y, w)
    for out in (out1, out2):
        assert rel_err(out["beams"], ref_beams) < BEAM_TOL


def test_corr_beam_float_dump_and_reset():
    shape = syn.BlockShape(C=8, F=1, NT=64, R=36, P=2, RPP=36, NP=2, B=4)
    rng = np.random.default_rng(5)
    ring = rng.integers(-127, 128, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    scales = np.ones(shape.scale_shape(), np.int16)
    p = lam.Pipeline(shape, lam.MODE_CORR_BEAM)
    p.set_weights(syn.random_weights(shape, 1)[0])
    p.run_host(ring, scales)
    f = p.dump_visibilities()
    ref = orc.correlate(_o(shape), orc.ingest(_o(shape), ring, scales, orc.WINDOW_CORR_BEAM))
    assert np.array_equal(f, ref.astype(np.float32))  # |V| < 2^24: exact in float
    assert p.blocks_accumulated() == 0
    again = p.dump_visibilities_i64()
    assert not again.any()
    p.close()


def test_scales_and_missing_packets():
    """Per-packet int16 scales (zeroed scale = missing packet,
    packet_formats.hpp:94-112) applied before correlation."""
    shape = syn.BlockShape(C=2, F=2, NT=16, R=8, P=2, RPP=4, NP=4, B=2)
    rng = np.random.default_rng(11)
    ring = rng.integers(-30, 31, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    scales = rng.integers(0, 4, shape.scale_shape()).astype(np.int16)
    scales[0, 2, 4:8] = 0
    p = lam.Pipeline(shape, lam.MODE_CORR_BEAM)
    w16, w = syn.random_weights(shape, 2)
    p.set_weights(w16)
    out = p.run_host(ring, scales)
    vis = p.dump_visibilities_i64()
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_CORR_BEAM)
    assert np.array_equal(vis, orc.correlate(s, y))
    assert rel_err(out["beams"], orc.beamform(s, y, w)) < BEAM_TOL


def test_range_error_is_reported():
    shape = SMALL
    ring, scales = syn.ring_from_fn(shape, lambda *a: (100, 0), lambda *a: 3)
    p = lam.Pipeline(shape, lam.MODE_CORR_BEAM)
    p.set_weights(syn.unity_weights(shape)[0])
    beams = p.beams_array()
    p.execute(ring, scales, beams_out=beams)
    with pytest.raises(lam.LamError) as ei:
        p.sync()
    assert ei.value.code == lam.LAM_ERR_RANGE
    p.close()


def test_range_error_fails_the_next_call_and_releases_once():
    """A block with |x s| > 127 fails the first call after it completes (no
    silent corruption in a feeder loop that never calls lam_sync); every
    host block handed to lam_execute is released exactly once, also when the
    call fails."""
    import time
    shape = SMALL
    bad_ring, bad_scales = syn.ring_from_fn(shape, lambda *a: (100, 0), lambda *a: 3)
    ring, scales = syn.ring_from_fn(shape, lambda *a: (2, -2))
    p = lam.Pipeline(shape, lam.MODE_CORR_BEAM)
    p.set_weights(syn.unity_weights(shape)[0])
    released = []
    cb = lam.HOSTFN(lambda user: released.append(int(user or 0)))
    p.execute(bad_ring, bad_scales, release_fn=cb, release_user=1)
    time.sleep(0.5)
    with pytest.raises(lam.LamError) as ei:
        p.execute(ring, scales, release_fn=cb, release_user=2)
    assert ei.value.code == lam.LAM_ERR_RANGE
    assert "block 0" in str(ei.value)
    p.dump_visibilities_i64()  # drops the corrupted block's visibilities
    # the error was reported once: the next block runs
    p.execute(ring, scales, release_fn=cb, release_user=3)
    # a call that fails its own checks still releases its block
    with pytest.raises(lam.LamError):
        p.execute(ring, scales, spectra_out=np.zeros(4, np.float32), release_fn=cb, release_user=4)
    p.sync()
    assert sorted(released) == [1, 2, 3, 4]
    # blocks 2 and 4 never reached the accumulator
    vis = p.dump_visibilities_i64()
    s = _o(shape)
    assert np.array_equal(vis, orc.correlate(s, orc.ingest(s, ring, scales, orc.WINDOW_CORR_BEAM)))
    p.close()


def test_graphs_replay_and_disable_switch(monkeypatch):
    """Blocks replay hipGraphs by default; SPATIAL_DISABLE_CUDA_GRAPH (any
    value, lambda_corr_beam_only_gpu_pipeline.hpp:561) forces eager launches
    with identical results."""
    shape = syn.BlockShape(C=2, F=1, NT=64, R=36, P=2, RPP=36, NP=2, B=4)
    rng = np.random.default_rng(3)
    ring = rng.integers(-127, 128, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    scales = np.ones(shape.scale_shape(), np.int16)
    w16 = syn.random_weights(shape, 1)[0]
    outs = []
    for disable in (False, True):
        if disable:
            monkeypatch.setenv("SPATIAL_DISABLE_CUDA_GRAPH", "")
        p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
        p.set_weights(w16)
        p.set_fixed_k(np.ones((shape.C, 2), np.int32))
        for _ in range(3):
            out = p.run_host(ring, scales)
        replays = p.graph_replays()
        p.close()
        assert (replays == 0) if disable else (replays >= 3)
        outs.append(out["beams"])
    assert np.array_equal(outs[0], outs[1])


def test_gains_and_fine_delays():
    """Gains g_i conj(g_j) on visibilities and folded into the beams; fine
    delays as the channel-centre phase exp(-2 pi i f_c tau)."""
    shape = syn.BlockShape(C=2, F=1, NT=32, R=8, P=2, RPP=8, NP=2, B=2)
    rng = np.random.default_rng(4)
    ring = rng.integers(-60, 61, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    scales = np.ones(shape.scale_shape(), np.int16)
    g = (rng.standard_normal((2, 2, 8)) + 1j * rng.standard_normal((2, 2, 8))).astype(np.complex64)
    tau = rng.uniform(-5, 5, 8).astype(np.float32)
    p = lam.Pipeline(shape, lam.MODE_CORR_BEAM)
    w16, w = syn.random_weights(shape, 6)
    p.set_weights(w16)
    p.set_gains(g)
    p.set_fine_delays(tau, 0.0, 781250.0, 64)
    out = p.run_host(ring, scales)
    vis = p.dump_visibilities()
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_CORR_BEAM)
    raw = orc.correlate(s, y)
    fc = (64 + np.arange(2)) * 781250.0
    ph = np.exp(-2j * np.pi * fc[:, None] * tau[None, :].astype(np.float64) * 1e-9)  # [C][R]
    ge = g.astype(np.complex128) * ph[:, None, :]  # effective per-(c,p,r) factor
    ref = np.zeros(raw.shape[:-1], complex)
    for j in range(8):
        for i in range(j + 1):
            v = raw[:, bl(i, j), :, :, 0] + 1j * raw[:, bl(i, j), :, :, 1]
            ref[:, bl(i, j)] = v * ge[:, :, i][:, :, None] * np.conj(ge[:, :, j])[:, None, :]
    got = vis[..., 0] + 1j * vis[..., 1]
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-5
    ref_beams = orc.beamform(s, y, w * ge[:, :, None, :])
    assert rel_err(out["beams"], ref_beams) < BEAM_TOL


# ---------------------------------------------------------- eigensolver
def _special_hermitian(n, rng):
    """Matrices that stress an eigensolver: degenerate spectra (the
    all-ones rank-1 matrix: eigenvalue n and n - 1 zeros, the reference's
    EigenvalueBasic shape), a diagonal one (already reduced), a complex
    tridiagonal one, the zero matrix, and a strong rank-2 interferer over noise
    whitened like a covariance."""
    out = [np.ones((n, n), complex), np.diag(rng.standard_normal(n)).astype(complex), np.zeros((n, n), complex)]
    T = np.diag(rng.standard_normal(n)).astype(complex)
    if n > 1:
        off = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
        T += np.diag(off, -1) + np.diag(off.conj(), 1)
    out.append(T)
    X = rng.standard_normal((n, 4 * n)) + 1j * rng.standard_normal((n, 4 * n))
    a = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    X += 30 * a @ (rng.standard_normal((2, 4 * n)) + 1j * rng.standard_normal((2, 4 * n)))
    C = X @ X.conj().T
    d = np.sqrt(np.diag(C).real)
    out.append(C / np.outer(d, d))
    return out


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 12, 24, 36, 64])
def test_eigh_batched_vs_oracle(n):
    """The batched Jacobi against the FP64 oracle Jacobi: eigenvalues,
    residuals and orthogonality, on random covariances and the special
    matrices above."""
    _log(f"eigh n={n}")
    rng = np.random.default_rng(n)
    nmat = 7
    As = []
    for _ in range(nmat):
        X = rng.standard_normal((n, 2 * n)) + 1j * rng.standard_normal((n, 2 * n))
        As.append(X @ X.conj().T / (2 * n))
    As += _special_hermitian(n, rng)
    nmat = len(As)
    As = np.array(As)
    buf = np.stack([As.transpose(0, 2, 1).real, As.transpose(0, 2, 1).imag], -1).astype(np.float32)
    dA = lam.DeviceArray.from_numpy(buf)
    dev = lam.DeviceArray((nmat, n), np.float32)
    dvec = lam.DeviceArray((nmat, n, n, 2), np.float32)
    dsw = lam.DeviceArray((nmat,), np.int32)
    lam.eigh_batched(dA, n, nmat, dev, dvec, dsw, max_sweeps=20, tol=1e-7)
    ev = dev.numpy()
    V = dvec.numpy()
    V = V[..., 0] + 1j * V[..., 1]  # [m][k][i]
    for m in range(nmat):
        ref_ev, ref_V, _ = orc.jacobi_eigh(As[m])
        scale = max(np.max(np.abs(ref_ev)), 1e-30)
        assert np.max(np.abs(ev[m] - ref_ev)) / scale < 2e-5, m
        U = V[m].T  # columns
        resid = np.linalg.norm(As[m] @ U - U * ev[m][None, :]) / max(np.linalg.norm(As[m]), 1e-30)
        assert resid < 1e-4, m
        assert np.allclose(U.conj().T @ U, np.eye(n), atol=1e-4), m


# ---------------------------------------------------------- adaptive path
def adaptive_run(shape, sample_fn, K=1, shrink=False, detect=False, delta=3.0, weights=None):
    ring, scales = syn.ring_from_fn(shape, sample_fn)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
    p.set_weights(syn.unity_weights(shape)[0] if weights is None else weights)
    p.set_fixed_k(np.full((shape.C, 2), K))
    p.set_nulling(detect=detect, delta=delta, shrink=shrink)
    p.set_eig_params(20, 1e-7)
    out = p.run_host(ring, scales)
    p.close()
    return out


def test_adaptive_null_rank_one():
    """test_adaptive_pipeline.cu:414-423, 491-511."""
    out = adaptive_run(ADAPT, lambda *a: (2, -2), K=1)
    b = out["beams"]
    assert np.all(b[0, :, 0] == 8 - 8j)
    assert np.sum(np.abs(b[0, :, 1]) ** 2) < 1.0
    assert np.all(np.abs(b[0, :, 1].real) <= 0.25) and np.all(np.abs(b[0, :, 1].imag) <= 0.25)
    ev = out["evals_out"][0]
    assert np.all(np.diff(ev, axis=-1) >= -1e-3) and np.all(ev >= -1e-3)


def test_adaptive_k0_tagged_layout():
    """test_adaptive_pipeline.cu:390-489, 539-567 (null and shrink)."""
    def tagged(c, f, k, t, r, p):
        return (t + 1, -(1 + 10 * p + 20 * c)) if r == 0 else (0, 0)
    for shape in (ADAPT, syn.BlockShape(C=2, F=1, NT=64, R=4, P=2, RPP=4, NP=1, B=1)):
        for shrink in (False, True):
            out = adaptive_run(shape, tagged, K=0, shrink=shrink)
            b = out["beams"]
            for c in range(shape.C):
                for p in range(2):
                    exp = np.arange(1, 65) - 1j * (1 + 10 * p + 20 * c)
                    assert np.array_equal(b[c, p, 0], exp) and np.array_equal(b[c, p, 1], exp)


def test_adaptive_shrink_vs_null_and_detection():
    """test_adaptive_pipeline.cu:622-674."""
    out = adaptive_run(ADAPT, lambda *a: (2, -2), K=1, shrink=True)
    assert np.sum(np.abs(out["beams"][0, :, 1]) ** 2) < 1.0

    def rank2(c, f, k, t, r, p):
        if t < 32:
            return (2, -2)
        return (4, 0) if r == 0 else (0, 0)
    pn = np.sum(np.abs(adaptive_run(ADAPT, rank2, K=1)["beams"][0, :, 1]) ** 2)
    ps = np.sum(np.abs(adaptive_run(ADAPT, rank2, K=1, shrink=True)["beams"][0, :, 1]) ** 2)
    assert ps > pn + 1e-6

    def multimode(c, f, k, t, r, p):
        if p == 1:
            return (0, 0)
        a = 20 if (t // 8) % 2 == 0 else -20
        b = 18 if (t // 4) % 2 == 0 else -18
        return (a + (1 if r % 2 == 0 else -1) * b, 0)
    out = adaptive_run(ADAPT, multimode, K=0, detect=True, delta=0.5)
    assert list(out["counts_out"][0]) == [2, 0]
    assert np.sum(np.abs(out["beams"][0, 0, 1]) ** 2) < 1.0
    assert np.sum(np.abs(out["beams"][0, 1, 1]) ** 2) == 0.0


@pytest.mark.parametrize("shrink", [False, True])
def test_adaptive_scene_vs_oracle(shrink):
    """Synthetic scene (2 CW tones, 36 antennas): eigenvalues, K, projected
    beams and the nulled subspace against the FP64 oracle."""
    shape = syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=4, B=4)
    _log(f"adaptive scene shrink={shrink}")
    scene = syn.make_scene(shape, seed=21, n_tones=2, inr_db=(10, 25))
    ring, scales = syn.make_block(shape, scene, seed=22)
    w16, w = syn.random_weights(shape, 23)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
    p.set_weights(w16)
    p.set_fixed_k(np.full((shape.C, 2), 2))
    p.set_nulling(shrink=shrink)
    p.set_eig_params(20, 1e-7)
    out = p.run_host(ring, scales)
    vis = p.dump_visibilities_i64()
    orth = p.eig_stats(with_orthogonality=True)[2]
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)
    ref = orc.adaptive_block(s, y, w, fixed_k=np.full((shape.C, 2), 2), shrink=shrink)
    # FP64 orthogonality check of the removed basis, max |U^H U - I|
    assert orth.max() < 1e-5
    if shrink:  # shrink mode returns the whitened vectors themselves: recompute on the CPU
        V = out["evecs_out"][..., 0].astype(np.float64) + 1j * out["evecs_out"][..., 1]
        U = V[:, :, -2:, :]
        G = np.einsum("cpai,cpbi->cpab", U.conj(), U) - np.eye(2)
        assert np.allclose(orth, np.abs(G).max(axis=(2, 3)), rtol=1e-3, atol=1e-10)
    assert np.array_equal(vis, ref["vis"])
    evg = out["evals_out"]
    assert np.max(np.abs(evg - ref["evals"]) / np.max(np.abs(ref["evals"]), axis=-1, keepdims=True)) < 1e-4
    assert list(out["counts_out"].ravel()) == list(ref["counts"].ravel())
    assert rel_err(out["beams"][:, :, :4], ref["beams2b"][:, :, :4]) < BEAM_TOL
    assert rel_err(out["beams"][:, :, 4:], ref["beams2b"][:, :, 4:]) < BEAM_TOL
    # nulled subspace: projector onto the K largest eigenvectors
    V = out["evecs_out"][..., 0] + 1j * out["evecs_out"][..., 1]
    for c in range(shape.C):
        for pp in range(2):
            Ug = V[c, pp, -2:].T
            Ur = ref["evecs"][c, pp, -2:].T
            assert np.max(np.abs(Ug @ Ug.conj().T - Ur @ Ur.conj().T)) < 1e-4


@pytest.mark.parametrize("R,K,detect,tones", [(36, 1, False, 1), (36, 2, False, 2), (36, 4, False, 2),
                                               (36, 0, True, 3), (12, 2, False, 2), (3, 1, False, 1),
                                               (64, 3, False, 3), (36, 10, False, 2)])
def test_adaptive_without_eigen_outputs(R, K, detect, tones):
    """A block that requests no eigenvalues or eigenvectors takes the
    tridiagonal solve (reduction, multisection, inverse iteration; more than
    8 removed modes: the Jacobi solver): K and the projected beams against
    the FP64 oracle, and against the same block with eigen outputs (the
    Jacobi solver).  An all-zero channel rides along."""
    shape = syn.BlockShape(C=4, F=1, NT=64, R=R, P=2, RPP=R, NP=4, B=4)
    _log(f"adaptive without eigen outputs R={R} K={K} detect={detect}")
    scene = syn.make_scene(shape, seed=31 + R + K, n_tones=tones, inr_db=(8, 30))
    ring, scales = syn.make_block(shape, scene, seed=32 + R)
    ring[3] = 0  # channel 3 silent: zero covariance
    w16, w = syn.random_weights(shape, 33)
    outs = []
    for eigen in (False, True):
        p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
        p.set_weights(w16)
        p.set_fixed_k(np.full((shape.C, 2), K))
        p.set_nulling(detect=detect, delta=3.0)
        p.set_eig_params(20, 1e-7)
        outs.append(p.run_host(ring, scales, eigen=eigen))
        outs[-1]["sweeps"] = p.eig_stats()[0]
        p.close()
    tri, jac = outs
    assert "evals_out" not in tri
    # the tridiagonal solve reports no sweeps; beyond 8 vectors the Jacobi solver ran
    assert (tri["sweeps"][:3] > 0).all() if K > 8 else (tri["sweeps"] == 0).all()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)
    ref = orc.adaptive_block(s, y, w, fixed_k=np.full((shape.C, 2), K), detect=detect, delta=3.0)
    assert list(tri["counts_out"].ravel()) == list(ref["counts"].ravel())
    assert list(tri["counts_out"].ravel()) == list(jac["counts_out"].ravel())
    B = shape.B
    for lo, hi in ((0, B), (B, 2 * B)):
        assert rel_err(tri["beams"][:, :, lo:hi], ref["beams2b"][:, :, lo:hi]) < BEAM_TOL
        assert rel_err(tri["beams"][:, :, lo:hi], jac["beams"][:, :, lo:hi]) < BEAM_TOL
    assert not tri["beams"][3].any()


def test_adaptive_results_independent_of_split():
    """The correlator's time split (the automatic rule, 1024 samples, one split)
    changes only how the int32 partial sums are cut: the int64 visibilities equal
    the oracle's for every split, and since the eigen stage sums the partials
    exactly and the 2B-beam launch takes its own split, the beams and the mode
    counts are bit-identical across splits."""
    shape = syn.BlockShape(C=8, F=1, NT=64, R=36, P=2, RPP=36, NP=64, B=4)
    scene = syn.make_scene(shape, seed=71, n_tones=2, inr_db=(8, 30))
    ring, scales = syn.make_block(shape, scene, seed=72)
    w16, _ = syn.random_weights(shape, 73)
    s = _o(shape)
    vis_ref = orc.correlate(s, orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE))
    outs = []
    for spw in (0, 1024, 8192):
        _log(f"adaptive split independence spw={spw}")
        p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
        p.set_weights(w16)
        p.set_fixed_k(np.full((shape.C, 2), 2))
        p.set_eig_params(20, 1e-7)
        p.set_tuning(spw)
        out = p.run_host(ring, scales, eigen=False)
        vis = p.dump_visibilities_i64()
        p.close()
        assert np.array_equal(vis, vis_ref)
        outs.append(out)
    for o in outs[1:]:
        assert np.array_equal(o["beams"], outs[0]["beams"])
        assert np.array_equal(o["counts_out"], outs[0]["counts_out"])


def test_adaptive_spectra_vs_oracle():
    """Post-beam spectra (register_fft_block product, adaptive 808-848): the
    GPU spectra of the GPU's own fp16 beams match the FP64 DFT oracle within
    1e-5 of each row's largest bin, and the FP64 oracle's beams within 1e-3
    (the reference transforms its FP32 beams; ours reads the fp16 output, so
    the fp16 rounding of the beams is the whole difference).  Two blocks in a
    row through the captured graphs, with three buffers."""
    shape = syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=4, B=4)
    _log("adaptive spectra scene")
    scene = syn.make_scene(shape, seed=31, n_tones=2, inr_db=(10, 25))
    w16, w = syn.random_weights(shape, 33)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE, num_buffers=3)
    p.set_weights(w16)
    p.set_fixed_k(np.full((shape.C, 2), 2))
    p.set_eig_params(20, 1e-7)
    p.set_spectra(True)
    s = _o(shape)
    for blk in range(2):
        ring, scales = syn.make_block(shape, scene, seed=40 + blk, time_offset=blk * shape.T)
        out = p.run_host(ring, scales)
        sg = out["spectra_out"].astype(np.float64)
        own = orc.beam_spectra(out["beams"].astype(np.complex128))
        rowmax = np.max(own, axis=-1, keepdims=True)
        assert np.max(np.abs(sg - own) / rowmax) < 1e-5
        y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)
        ref = orc.adaptive_block(s, y, w, fixed_k=np.full((shape.C, 2), 2))
        fp64 = orc.beam_spectra(ref["beams2b"])
        assert rel_err(sg, fp64) < BEAM_TOL
    p.close()


def test_spectra_many_rows_partial_tiles():
    """5120 (channel, pol, beam) rows of 24 packets: more rows than the
    persistent grid has waves, so each wave walks several rows with the next
    tile prefetched across row boundaries, and every row ends on a partial
    16-packet tile.  Spectra of the GPU's own fp16 beams within 1e-5 of each
    row's largest bin."""
    shape = syn.BlockShape(C=80, F=1, NT=64, R=36, P=2, RPP=36, NP=24, B=16)
    _log("spectra many rows")
    scene = syn.make_scene(shape, seed=71, n_tones=1)
    ring, scales = syn.make_block(shape, scene, seed=72)
    w16, _ = syn.random_weights(shape, 73)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
    p.set_weights(w16)
    p.set_fixed_k(np.ones((shape.C, 2)))
    p.set_spectra(True)
    out = p.run_host(ring, scales)
    p.close()
    own = orc.beam_spectra(out["beams"].astype(np.complex128))
    rowmax = np.max(own, axis=-1, keepdims=True)
    assert np.max(np.abs(out["spectra_out"] - own) / rowmax) < 1e-5


def test_adaptive_spectra_tone_bins():
    """A tone at DFT bin k on every receiver peaks at output bin k + 27
    (k < 27) or k - 37 (k >= 37) in both beams with K = 0; bins 27..36 are
    removed, so such a tone leaves only leakage from int8 rounding."""
    for k in (0, 5, 20, 40, 63, 30):
        def tone(c, f, pk, t, r, pol, k=k):
            a = 2 * np.pi * k * t / 64
            return (int(round(40 * np.cos(a))), int(round(40 * np.sin(a))))
        ring, scales = syn.ring_from_fn(ADAPT, tone)
        p = lam.Pipeline(ADAPT, lam.MODE_ADAPTIVE)
        p.set_weights(syn.unity_weights(ADAPT)[0])
        p.set_fixed_k(np.zeros((1, 2)))
        p.set_spectra(True)
        out = p.run_host(ring, scales)
        p.close()
        sp = out["spectra_out"]
        own = orc.beam_spectra(out["beams"].astype(np.complex128))
        assert np.max(np.abs(sp - own)) <= 1e-5 * np.max(own)
        for pol in range(2):
            for b in range(2):
                row = sp[0, pol, b]
                if 27 <= k <= 36:
                    assert np.max(row) < 0.02 * 4 * 40 * 64
                else:
                    assert int(np.argmax(row)) == (k + 27 if k < 27 else k - 37)
                    assert abs(row.max() - 4 * 40 * 64) < 0.01 * 4 * 40 * 64


def test_spectra_argument_errors():
    """Spectra exist only in adaptive mode; asking for them when off is an error."""
    p = lam.Pipeline(SMALL, lam.MODE_CORR_BEAM)
    with pytest.raises(lam.LamError):
        p.set_spectra(True)
    p.close()
    p = lam.Pipeline(ADAPT, lam.MODE_ADAPTIVE)
    ring, scales = syn.ring_from_fn(ADAPT, lambda *a: (1, 0))
    with pytest.raises(lam.LamError):
        p.execute(ring, scales, spectra_out=p.spectra_array())
    p.set_spectra(True)
    p.set_spectra(False)
    p.close()


def test_host_ring_overlapped_ingest():
    """Overlapped ingest (ProcessorState -> pinned buffers -> async H2D,
    spatial.hpp:718-770, 1182-1299): 9 blocks through 3 pinned buffers, each
    buffer refilled with the next block as soon as the pipeline's release
    callback returns it.  The release comes after the block's H2D copy, so
    the accumulated visibilities equal the oracle's sum over the 9 blocks
    exactly; a release before the copy had finished would mix blocks."""
    shape = syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=4, B=4, blocks_to_accumulate=100)
    _log("host ring ingest")
    scene = syn.make_scene(shape, seed=51, n_tones=1)
    blocks = [syn.make_block(shape, scene, seed=60 + b, time_offset=b * shape.T) for b in range(9)]
    p = lam.Pipeline(shape, lam.MODE_CORR_BEAM, num_buffers=3)
    p.set_weights(syn.unity_weights(shape)[0])
    ring = lam.HostRing(p, 3)
    for ring_b, scales_b in blocks:
        i = ring.acquire()
        ring.samples[i].array[...] = ring_b
        ring.scales[i].array[...] = scales_b
        ring.submit(i)
    p.sync()
    vis = p.dump_visibilities_i64()
    assert ring.released == 9
    ring.close()
    p.close()
    s = _o(shape)
    ref = sum(orc.correlate(s, orc.ingest(s, r, sc, orc.WINDOW_CORR_BEAM)) for r, sc in blocks)
    assert np.array_equal(vis, ref)


def test_adaptive_delays_and_permutation():
    """apply_delays (spatial.cuh:1420-1506) and reorder_streams
    (spatial.cuh:593-649) folded into the voltage loads; K=0 beams = W x."""
    shape = syn.BlockShape(C=2, F=2, NT=16, R=8, P=2, RPP=4, NP=3, B=2)
    rng = np.random.default_rng(9)
    ring = rng.integers(-50, 51, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    scales = np.ones(shape.scale_shape(), np.int16)
    delays = np.array([3, -5], np.int32)
    recv = np.array([1, 1, 0, 2, -1, 3, 4, 5, 6, 6, 7, 7, 5, 4, 0, 0], np.int32)
    pol = np.array([1, 0, 0, 1, 0, 0, 1, 1, 0, 1, 1, 0, 0, 1, 1, 0], np.int32)
    w16, w = syn.random_weights(shape, 12)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
    p.set_weights(w16)
    p.set_fixed_k(np.zeros((2, 2)))
    p.set_subpacket_delays(delays)
    p.set_stream_permutation(recv, pol)
    out = p.run_host(ring, scales)
    vis = p.dump_visibilities_i64()
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE, delays=delays, recv_perm=recv, pol_perm=pol)
    assert np.array_equal(vis, orc.correlate(s, y))
    ref = orc.beamform(s, y, w)
    assert rel_err(out["beams"][:, :, :2], ref) < BEAM_TOL
    assert rel_err(out["beams"][:, :, 2:], ref) < BEAM_TOL


@pytest.mark.parametrize("num_buffers", [1, 3])
def test_autotune_leaves_results_exact(num_buffers):
    """The autotuner times launch shapes on the caller's blocks; afterwards the
    pipeline's visibilities hold only the blocks executed after it (bit-exact)
    and the beams stay within tolerance.  36 receivers with a partial last
    chunk and scales 0/1/2 exercise the packed-row loader.  With three buffers
    the split size is chosen by streaming time (overlapping blocks)."""
    shape = syn.BlockShape(C=6, F=1, NT=64, R=36, P=2, RPP=36, NP=5, B=4)
    rng = np.random.default_rng(11)
    ring = rng.integers(-60, 61, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    scales = rng.choice(np.array([0, 1, 1, 2], np.int16), shape.scale_shape())
    w16, w = syn.random_weights(shape, seed=4)
    p = lam.Pipeline(shape, lam.MODE_CORR_BEAM, num_buffers=num_buffers)
    p.set_weights(w16)
    d_ring = lam.DeviceArray.from_numpy(ring)
    d_scales = lam.DeviceArray.from_numpy(scales)
    res = p.autotune([d_ring, d_ring], [d_scales, d_scales], iters=1)
    _log(f"autotune {res}")
    assert res["samples_per_wg"] in (0, 1024, 2048, 3072, 4096, 8192)
    out = p.run_host(ring, scales)
    vis64 = p.dump_visibilities_i64()
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_CORR_BEAM)
    assert np.array_equal(vis64, orc.correlate(s, y))
    assert rel_err(out["beams"], orc.beamform(s, y, w)) < BEAM_TOL


def test_autotune_adaptive_warm_start():
    shape = syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=8, B=4)
    scene = syn.make_scene(shape, seed=21, n_tones=2, amp=30.0)
    blocks = [syn.make_block(shape, scene, seed=22 + i) for i in range(2)]
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
    p.set_weights(syn.random_weights(shape, seed=5)[0])
    p.set_fixed_k(np.full((shape.C, 2), 2, np.int32))
    p.learn_init(8, 256, seed=2)
    p.set_eig_method(lam.EIG_PREVIOUS, refine_sweeps=2)
    dev = [(lam.DeviceArray.from_numpy(r), lam.DeviceArray.from_numpy(s)) for r, s in blocks]
    res = p.autotune([d[0] for d in dev], [d[1] for d in dev], iters=1)
    _log(f"autotune {res}")
    assert res["plain_threads"] in (64, 128, 256) and res["refine_steps"] in (1, 2, 3)
    ring, scales = blocks[0]
    p.run_host(ring, scales)
    vis64 = p.dump_visibilities_i64()
    p.close()
    s = _o(shape)
    assert np.array_equal(vis64, orc.correlate(s, orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)))


@pytest.mark.parametrize("num_buffers", [1, 3])
def test_autotune_does_not_capture_training_blocks(num_buffers):
    """A pending lam_learn_capture records only the caller's blocks: the
    tuning blocks (plain solver, many passes) are not training examples."""
    shape = syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=8, B=4)
    scene = syn.make_scene(shape, seed=31, n_tones=1, amp=30.0)
    ring, scales = syn.make_block(shape, scene, seed=32)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE, num_buffers=num_buffers)
    p.set_weights(syn.random_weights(shape, seed=5)[0])
    p.set_fixed_k(np.ones((shape.C, 2), np.int32))
    p.learn_init(4, 64, seed=3)
    p.learn_capture(5)
    p.run_host(ring, scales)
    assert p.learn_captured() == 1
    d_ring, d_scales = lam.DeviceArray.from_numpy(ring), lam.DeviceArray.from_numpy(scales)
    p.autotune([d_ring], [d_scales], iters=1)
    assert p.learn_captured() == 1
    p.run_host(ring, scales)
    assert p.learn_captured() == 2
    p.close()


def subarray_beams_oracle(shape, vis_c, y_c, w_cp, member, pol, K):
    """Beams of one subarray group through the oracle's RFI-null chain
    (rfi_null_cp on the group's packed-triangular visibilities, then W_g' x_g)."""
    ant = np.flatnonzero(member)
    n = len(ant)
    sub = np.zeros((n * (n + 1) // 2, 2, 2, 2), np.int64)
    for j in range(n):
        for i in range(j + 1):
            sub[j * (j + 1) // 2 + i] = vis_c[bl(ant[i], ant[j])]
    W2, _, _, _ = orc.rfi_null_cp(n, 2, sub, pol, w_cp[:, ant], fixed_k=K)
    x = y_c[:, ant, pol, 0].astype(np.float64) + 1j * y_c[:, ant, pol, 1]  # [T][n]
    return W2[shape.B:] @ x.T


def test_subarrays_uneven_sizes_persistent():
    """Subarray groups of 12/24/36 antennas: every (channel, pol, group)
    eigenproblem in one persistent launch, against LAPACK on the oracle's
    principal submatrices; null projectors against I - U_K U_K^H."""
    shape = syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=8, B=4)
    scene = syn.make_scene(shape, seed=51, n_tones=3, amp=30.0)
    ring, scales = syn.make_block(shape, scene, seed=52)
    rng = np.random.default_rng(53)
    G = 2
    mask = np.zeros((shape.C, G, shape.R), np.uint8)
    for c in range(shape.C):
        for g in range(G):
            n = rng.choice([12, 24, 36])
            mask[c, g, rng.choice(shape.R, n, replace=False)] = 1
    k = rng.integers(1, 4, (shape.C, 2, G)).astype(np.int32)
    w16, w = syn.random_weights(shape, seed=54)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
    p.set_weights(w16)
    p.set_fixed_k(np.ones((shape.C, 2), np.int32))
    p.set_subarrays(mask, k)
    out = p.run_host(ring, scales)
    res = p.subarray_results()
    p.close()
    s = _o(shape)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)
    vis = orc.correlate(s, y)
    # the groups' mitigated beams (output beams 2B + gB + b) against the oracle's
    # RFI-null chain on each group's own visibilities
    assert out["beams"].shape[2] == (2 + G) * shape.B
    ref2b = orc.adaptive_block(s, y, w, fixed_k=np.ones((shape.C, 2)))["beams2b"]
    assert rel_err(out["beams"][:, :, :2 * shape.B], ref2b) < BEAM_TOL
    for c in range(shape.C):
        for pol in range(2):
            for g in range(G):
                ref = subarray_beams_oracle(shape, vis[c], y[c], w[c, pol], mask[c, g], pol, k[c, pol, g])
                got = out["beams"][c, pol, (2 + g) * shape.B:(3 + g) * shape.B]
                assert rel_err(got, ref) < BEAM_TOL, (c, pol, g)
    for c in range(shape.C):
        for pol in range(2):
            A = orc.covariance(shape.R, 2, vis[c], pol)
            for g in range(G):
                q = (c * 2 + pol) * G + g
                ant = np.flatnonzero(mask[c, g])
                n = len(ant)
                assert res["sizes"][q] == n and res["counts"][q] == k[c, pol, g]
                sub = A[np.ix_(ant, ant)]
                dd = np.sqrt(np.maximum(sub.diagonal().real, 1e-10))
                w, V = np.linalg.eigh(sub / np.outer(dd, dd))
                assert np.max(np.abs(res["evals"][q, :n] - w)) < 1e-4 * np.max(np.abs(w))
                K = k[c, pol, g]
                U = dd[:, None] * V[:, n - K:]
                U /= np.linalg.norm(U, axis=0)
                Pref = np.eye(n) - U @ U.conj().T
                Pg = res["proj"][q, :n * n].reshape(n, n).T  # column-major
                assert np.max(np.abs(Pg - Pref)) < 1e-3, (c, pol, g, n)
                # I - sum u_k u_k^H with unwhitened unit vectors (buildIdentityMinusProjection,
                # spatial.cuh:993-1034): Hermitian, a projector only for K = 1
                assert np.max(np.abs(Pg - Pg.conj().T)) < 1e-5
                if K == 1:
                    assert np.max(np.abs(Pg @ Pg - Pg)) < 1e-4


def test_overlapped_buffers_match_serial():
    """Three buffers with graph replay (blocks overlap: the correlator and the
    2B beams of neighbouring blocks run beside each other's eigen stages, only
    the eigen stages are chained) against one buffer run eagerly: the same
    beams, eigenpairs, subarray results and accumulated visibilities, bit for
    bit, over a warm-started block sequence (pipeline/common.hpp:388-396,
    lambda_adaptive_beamformed_spectra_pipeline.hpp:820-880)."""
    shape = syn.BlockShape(C=8, F=1, NT=64, R=36, P=2, RPP=36, NP=16, B=4)
    nblk = 7
    blocks = []
    for i in range(nblk):
        scene = syn.make_scene(shape, seed=61 + i, n_tones=2, amp=30.0)
        blocks.append(syn.make_block(shape, scene, seed=161 + i))
    dev = [(lam.DeviceArray.from_numpy(r), lam.DeviceArray.from_numpy(s)) for r, s in blocks]
    rng = np.random.default_rng(62)
    G = 2
    mask = np.zeros((shape.C, G, shape.R), np.uint8)
    for c in range(shape.C):
        for g in range(G):
            mask[c, g, rng.choice(shape.R, rng.choice([12, 24, 36]), replace=False)] = 1
    sub_k = rng.integers(1, 3, (shape.C, 2, G)).astype(np.int32)
    w16 = syn.random_weights(shape, seed=63)[0]

    def run(nbuf, graphs):
        _log(f"overlap run nbuf={nbuf} graphs={graphs}")
        p = lam.Pipeline(shape, lam.MODE_ADAPTIVE, num_buffers=nbuf)
        p.set_weights(w16)
        p.set_fixed_k(np.full((shape.C, 2), 2, np.int32))
        p.set_subarrays(mask, sub_k)
        p.learn_init(8, 256, seed=4)
        p.set_graphs(graphs)
        outs = []
        for i in range(nblk):
            if i == 1:
                p.set_eig_method(lam.EIG_PREVIOUS, refine_sweeps=2)
            beams = p.beams_array()
            ev = np.zeros((shape.C, 2, shape.R), np.float32)
            vec = np.zeros((shape.C, 2, shape.R, shape.R, 2), np.float32)
            cnt = np.zeros((shape.C, 2), np.int32)
            p.execute(dev[i][0], dev[i][1], beams_out=beams, evals_out=ev, evecs_out=vec,
                      counts_out=cnt, on_device=True)
            outs.append((beams, ev, vec, cnt))
        p.sync()
        sub = p.subarray_results()
        vis = p.dump_visibilities_i64()
        p.close()
        return outs, sub, vis

    o1, s1, v1 = run(1, False)
    o3, s3, v3 = run(3, True)
    assert np.array_equal(v1, v3)
    for i in range(nblk):
        for a, b in zip(o1[i], o3[i]):
            assert np.array_equal(a, b), i
    for key in s1:
        assert np.array_equal(s1[key], s3[key]), key
    s = _o(shape)
    ref = sum(orc.correlate(s, orc.ingest(s, r, sc, orc.WINDOW_ADAPTIVE)) for r, sc in blocks)
    assert np.array_equal(v3, ref)


def test_warm_start_rejection_reruns_plain_solver():
    """A warm start that fails its leakage test (stale proposal, tolerance 0)
    re-solves the problem with the plain solver inside the same launch: the
    outputs equal a plain-only run bit for bit and the fallback flags are set."""
    shape = syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=8, B=4)
    blocks = [syn.make_block(shape, syn.make_scene(shape, seed=71 + i, n_tones=2, amp=30.0), seed=81 + i)
              for i in range(2)]
    w16 = syn.random_weights(shape, seed=73)[0]

    def run(warm):
        p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
        p.set_weights(w16)
        p.set_fixed_k(np.full((shape.C, 2), 2, np.int32))
        p.learn_init(8, 256, seed=1)
        p.run_host(*blocks[0])
        if warm:
            p.set_eig_method(lam.EIG_PREVIOUS, refine_sweeps=1, residual_tol=0.0)
        out = p.run_host(*blocks[1])
        sw, fb = p.eig_stats()
        p.close()
        return out, sw, fb

    ref, sw0, fb0 = run(False)
    got, sw1, fb1 = run(True)
    assert fb0.sum() == 0 and np.all(fb1 == 1)
    assert np.all(sw1 == sw0 + 1)  # the refinement step is counted with the plain sweeps
    for key in ("beams", "evals_out", "evecs_out", "counts_out"):
        assert np.array_equal(got[key], ref[key]), key
