"""GPU tests at the benchmark configurations' own sizes (384 channels x 2 pols
x 36 antennas in the 64-receiver layout, 8192 samples per block), against
the CPU oracle and LAPACK.

Bars: visibilities bit-exact (integers); beams within 1e-3 normwise relative
error of the FP64 oracle (FP16 hi+lo weights); eigenvalues of the nulled
subspace within 1e-5 relative of LAPACK (numpy eigvalsh) on the FP64
whitened covariance of the block's exact visibilities, on every (channel,
pol); the other eigenvalues within 1e-4 of each matrix's largest."""
import sys

import numpy as np
import pytest

import lam
import lam_synth as syn
import oracle as orc

pytestmark = pytest.mark.gpu

C, R, NP, B = 384, 36, 128, 16
SHAPE = syn.BlockShape(C=C, F=1, NT=64, R=R, P=2, RPP=R, NP=NP, B=B)
OSHAPE = orc.Shape(C=C, F=1, NT=64, R=R, P=2, RPP=R, NP=NP, B=B)
EIG_TOL = 1e-5
BEAM_TOL = 1e-3


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _synth(n_src, inr_db, seed, block=0, drift=0.0):
    sc = syn.DriftScenario(C, R, seed=seed, n_blocks=block + 1, n_src=(n_src, n_src), inr_db=inr_db,
                           drift=drift, p_switch=0.0, max_sources=n_src)
    tones, st, gain, _, _ = sc.block(block)
    dr = lam.DeviceArray(SHAPE.ring_shape(), np.int8)
    ds = lam.DeviceArray(SHAPE.scale_shape(), np.int16)
    dt, dst, dg = (lam.DeviceArray.from_numpy(a) for a in (tones, st, gain))
    lam.synth_block(SHAPE, seed, block, 16.0, dt, dst, dg, n_src, dr, ds)
    return dr, ds, tones


def _whitened(vis):
    """FP64 whitened XX / YY covariances [C][2][R][R] of exact visibilities."""
    Vc = vis[..., 0].astype(np.float64) + 1j * vis[..., 1].astype(np.float64)
    ii, jj = np.triu_indices(R)
    bl = jj * (jj + 1) // 2 + ii
    A = np.zeros((C, 2, R, R), np.complex128)
    for p in range(2):
        A[:, p, ii, jj] = Vc[:, bl, p, p]
        A[:, p, jj, ii] = np.conj(Vc[:, bl, p, p])
    d = np.sqrt(np.maximum(np.einsum("cpii->cpi", A).real, 1e-10))
    return A / (d[..., :, None] * d[..., None, :])


def _run(dr, ds, w16, K, wprec=lam.WPREC_F16_HILO):
    p = lam.Pipeline(SHAPE, lam.MODE_ADAPTIVE)
    p.set_weights(w16)
    p.set_fixed_k(np.full((C, 2), K, np.int32))
    p.set_eig_params(12, 1e-6)
    p.set_weight_precision(wprec)
    beams = p.beams_array()
    evals = np.zeros((C, 2, R), np.float32)
    evecs = np.zeros((C, 2, R, R, 2), np.float32)
    p.execute(dr, ds, beams_out=beams, evals_out=evals, evecs_out=evecs, on_device=True)
    p.sync()
    vis = p.dump_visibilities_i64()
    orth = p.eig_stats(with_orthogonality=True)[2]
    p.close()
    return beams, evals, evecs, vis, orth


def _complex(beams):
    return beams[..., 0].astype(np.float32) + 1j * beams[..., 1].astype(np.float32)


def test_cfg3_nulled_eigenvalues_every_channel():
    """3 interferers per channel at INR 0-30 dB, rank-3 null, each weight
    precision: the 3 removed eigenvalues of all 768 covariances within 1e-5
    of LAPACK and of the oracle's FP64 Jacobi; null depth from the beams."""
    dr, ds, tones = _synth(3, (0.0, 30.0), seed=20261203 + 3)
    w16, _ = syn.random_weights(SHAPE, seed=20261203 + 4)
    T = SHAPE.T
    ph = np.exp(-2j * np.pi * tones[..., 1].astype(np.float64)[..., None] * np.arange(T))  # [C][S][T]
    ref = None
    for name, mode in (("f16_hilo", lam.WPREC_F16_HILO), ("f16", lam.WPREC_F16), ("bf16", lam.WPREC_BF16)):
        _log(f"cfg3 {name}")
        beams, evals, _, vis, orth = _run(dr, ds, w16, 3, mode)
        if ref is None:
            Aw = _whitened(vis)
            ref = np.linalg.eigvalsh(Aw)
            # the oracle's FP64 Jacobi on a sample of the matrices agrees with LAPACK
            for c in range(0, C, 48):
                ev_o, _, _ = orc.jacobi_eigh(Aw[c, 1])
                assert np.max(np.abs(ev_o - ref[c, 1])) <= 1e-12 * ref[c, 1].max()
        top = np.abs(evals[..., -3:] - ref[..., -3:]) / np.abs(ref[..., -3:])
        rest = np.abs(evals[..., :-3] - ref[..., :-3]) / ref[..., -1:]
        depth = np.zeros((C, 2))
        for c0 in range(0, C, 32):  # known tones' amplitudes in the original and the nulled beams
            amp = np.einsum("cpbt,cst->cpbs", _complex(beams[c0:c0 + 32]), ph[c0:c0 + 32].astype(np.complex64)) / T
            orig = np.sum(np.abs(amp[:, :, :B]) ** 2, axis=(2, 3))
            mitig = np.sum(np.abs(amp[:, :, B:]) ** 2, axis=(2, 3))
            depth[c0:c0 + 32] = 10 * np.log10(np.maximum(mitig, 1e-30) / orig)
        _log(f"  top-3 eigenvalue rel err max {top.max():.2e}, others {rest.max():.2e}, "
             f"null depth median {np.median(depth):.1f} dB, orthogonality {orth.max():.1e}")
        assert top.max() <= EIG_TOL, (name, top.max(), np.unravel_index(np.argmax(top), top.shape))
        assert rest.max() <= 1e-4
        assert orth.max() < 1e-5
        assert np.median(depth) < -40.0


def test_cfg2_full_band_exact_and_sampled_beams():
    """cfg2: one CW tone per channel, 16 -> 32 beams, rank-1 null, visibilities
    accumulated over 4 blocks: visibilities bit-exact for every channel (one
    block) and for 16 sampled channels (4 blocks), beams and eigenvalues
    against the oracle on the sampled channels."""
    w16, w = syn.random_weights(SHAPE, seed=20261203 + 2)
    blocks = [_synth(1, (10.0, 30.0), seed=20261203 + 2, block=b, drift=0.01) for b in range(4)]
    p = lam.Pipeline(SHAPE, lam.MODE_ADAPTIVE, num_buffers=3)
    p.set_weights(w16)
    p.set_fixed_k(np.ones((C, 2), np.int32))
    p.set_eig_params(12, 1e-6)
    beams = p.beams_array()
    evals = np.zeros((C, 2, R), np.float32)
    p.execute(blocks[0][0], blocks[0][1], beams_out=beams, evals_out=evals, on_device=True)
    p.sync()
    vis1 = p.dump_visibilities_i64()
    for dr, ds, _ in blocks:
        p.execute(dr, ds, on_device=True)
    p.sync()
    vis4 = p.dump_visibilities_i64()
    p.close()
    rings = [(dr.numpy(), ds.numpy()) for dr, ds, _ in blocks]
    # every channel, block 0, in chunks of 48 channels
    for c0 in range(0, C, 48):
        s = orc.Shape(C=48, F=1, NT=64, R=R, P=2, RPP=R, NP=NP, B=B)
        yc = orc.ingest(s, rings[0][0][c0:c0 + 48], rings[0][1][c0:c0 + 48], orc.WINDOW_ADAPTIVE)
        assert np.array_equal(vis1[c0:c0 + 48], orc.correlate(s, yc)), c0
    _log("cfg2 visibilities of all 384 channels bit-exact")
    chans = np.arange(0, C, C // 16)
    s = orc.Shape(C=len(chans), F=1, NT=64, R=R, P=2, RPP=R, NP=NP, B=B)
    acc = 0
    for ring, scales in rings:
        acc = acc + orc.correlate(s, orc.ingest(s, ring[chans], scales[chans], orc.WINDOW_ADAPTIVE))
    assert np.array_equal(vis4[chans], acc)
    yb = _complex(beams[chans]).astype(np.complex128)
    y0 = orc.ingest(s, rings[0][0][chans], rings[0][1][chans], orc.WINDOW_ADAPTIVE)
    refb = orc.adaptive_block(s, y0, w[chans], fixed_k=np.ones((len(chans), 2)))
    eb = np.max(np.abs(yb - refb["beams2b"])) / np.max(np.abs(refb["beams2b"]))
    top = np.abs(evals[chans, :, -1] - refb["evals"][..., -1]) / np.abs(refb["evals"][..., -1])
    _log(f"cfg2 sampled: beam rel err {eb:.2e}, top eigenvalue rel err {top.max():.2e}")
    assert eb < BEAM_TOL
    assert top.max() <= EIG_TOL


def test_cfg4_one_queue_uneven_subarrays():
    """cfg4: 384 ch x 2 pol x 2 subarray groupings of 12/24/36 antennas (1536
    subarray problems) plus the 768 full-array problems of the block in one
    persistent launch, 1-3 tones at INR 10-30 dB: eigenvalues and projectors
    of a sample of problems against LAPACK, the groups' beams against the
    oracle on sampled channels, and the same results bit for bit from a
    7-workgroup grid (every workgroup takes ~330 problems from the queue)."""
    G = 2
    seed = 20261203 + 4
    sc = syn.DriftScenario(C, R, seed=seed, n_blocks=1, n_src=(1, 3), inr_db=(10.0, 30.0), drift=0.0,
                           p_switch=0.0, max_sources=3)
    tones, st, gain, _, act = sc.block(0)
    dr = lam.DeviceArray(SHAPE.ring_shape(), np.int8)
    ds = lam.DeviceArray(SHAPE.scale_shape(), np.int16)
    dt, dst, dg = (lam.DeviceArray.from_numpy(a) for a in (tones, st, gain))
    lam.synth_block(SHAPE, seed, 0, 16.0, dt, dst, dg, 3, dr, ds)
    rng = np.random.default_rng(seed)
    mask = np.zeros((C, G, R), np.uint8)
    sizes = rng.choice([12, 24, 36], (C, G))
    for c in range(C):
        for g in range(G):
            mask[c, g, rng.choice(R, sizes[c, g], replace=False)] = 1
    nk = act.sum(1).astype(np.int32)
    k = np.repeat(nk[:, None, None], 2, axis=1).repeat(G, axis=2)
    w16, w = syn.random_weights(SHAPE, seed=seed + 1)
    results = []
    for wgs in (0, 7):
        _log(f"cfg4 persistent grid {wgs or 'auto'}")
        p = lam.Pipeline(SHAPE, lam.MODE_ADAPTIVE)
        p.set_weights(w16)
        p.set_fixed_k(np.repeat(nk[:, None], 2, axis=1))
        p.set_subarrays(mask, k)
        p.set_persistent_workgroups(wgs)
        beams = p.beams_array()
        evals = np.zeros((C, 2, R), np.float32)
        p.execute(dr, ds, beams_out=beams, evals_out=evals, on_device=True)
        p.sync()
        res = p.subarray_results()
        vis = p.dump_visibilities_i64()
        p.close()
        results.append((beams, evals, res))
    (b0, e0, r0), (b1, e1, r1) = results
    assert np.array_equal(b0, b1) and np.array_equal(e0, e1)
    for key in ("evals", "evecs", "proj", "counts", "sizes", "sweeps"):
        assert np.array_equal(r0[key], r1[key]), key
    assert sorted(np.unique(r0["sizes"])) == [12, 24, 36]
    # LAPACK on a sample of the subarray problems (principal submatrices, whitened)
    Vc = vis[..., 0].astype(np.float64) + 1j * vis[..., 1].astype(np.float64)
    ii, jj = np.triu_indices(R)
    blv = jj * (jj + 1) // 2 + ii
    for q in range(0, C * 2 * G, 24):
        g, cp = q % G, q // G
        c, pol = cp // 2, cp % 2
        A = np.zeros((R, R), np.complex128)
        A[ii, jj] = Vc[c, blv, pol, pol]
        A[jj, ii] = np.conj(Vc[c, blv, pol, pol])
        ant = np.flatnonzero(mask[c, g])
        n = len(ant)
        sub = A[np.ix_(ant, ant)]
        dd = np.sqrt(np.maximum(sub.diagonal().real, 1e-10))
        lam_ref, V = np.linalg.eigh(sub / np.outer(dd, dd))
        Kq = k[c, pol, g]
        ev = r0["evals"][q, :n]
        assert np.max(np.abs(ev - lam_ref)) <= 1e-4 * lam_ref.max(), q
        if Kq:
            assert np.max(np.abs(ev[-Kq:] - lam_ref[-Kq:]) / np.abs(lam_ref[-Kq:])) <= EIG_TOL, q
        U = dd[:, None] * V[:, n - Kq:]
        U /= np.linalg.norm(U, axis=0)
        Pref = np.eye(n) - U @ U.conj().T
        Pg = r0["proj"][q, :n * n].reshape(n, n).T
        assert np.max(np.abs(Pg - Pref)) < 1e-3, q
    # the groups' beams on sampled channels against the oracle
    from test_gpu_parity import subarray_beams_oracle
    chans = np.arange(0, C, 48)
    s = orc.Shape(C=len(chans), F=1, NT=64, R=R, P=2, RPP=R, NP=NP, B=B)
    y = orc.ingest(s, dr.numpy()[chans], ds.numpy()[chans], orc.WINDOW_ADAPTIVE)
    for ci, c in enumerate(chans):
        for pol in range(2):
            for g in range(G):
                ref = subarray_beams_oracle(SHAPE, vis[c], y[ci], w[c, pol], mask[c, g], pol, k[c, pol, g])
                got = _complex(b0[c, pol, (2 + g) * B:(3 + g) * B]).astype(np.complex128)
                err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
                assert err < BEAM_TOL, (c, pol, g, err)
    _log("cfg4: 2304 problems in one launch; grid-independent; sample vs LAPACK and the oracle")


def test_cfg5_learned_previous_plain_384ch():
    """The learned warm start at its configuration's size: 384 channels x 2
    pols x 36 antennas, drifting switching interferers (1-4 per channel, INR
    5-30 dB, 0.01 rad/block, on/off p = 0.1), rank-4 null.  The plain
    solver's labels of 20 blocks train the network on the GPU (8 epochs; the
    loss falls), then 6 further blocks run with plain Jacobi, the previous
    block's vectors and the learned proposal.  Per method: the eigen-stage
    kernel time and the per-problem null depth against plain Jacobi's on the
    same blocks.  Bars: both warm starts accept nearly every problem, keep the
    mean null depth within 0.5 dB of plain Jacobi's, and run the eigen stage
    faster than plain Jacobi; the p99 of the per-problem |difference| is
    printed (it is dominated by the weakest sources, whose null depth sits at
    the noise floor)."""
    KW, K, NTR, NEV = 8, 4, 20, 6
    sc = syn.DriftScenario(C, R, seed=5, n_blocks=NTR + NEV)
    w16, W = syn.random_weights(SHAPE, seed=6)
    dr = lam.DeviceArray(SHAPE.ring_shape(), np.int8)
    ds = lam.DeviceArray(SHAPE.scale_shape(), np.int16)
    dt = lam.DeviceArray((C, 4, 6), np.float32)
    dst = lam.DeviceArray((C, 4, R, 2), np.float32)
    dg = lam.DeviceArray((C,), np.float32)

    def gen(b):
        tones, st, gain, stc, act = sc.block(b)
        dt.upload(tones)
        dst.upload(st)
        dg.upload(gain)
        lam.synth_block(SHAPE, 5, b, 16.0, dt, dst, dg, 4, dr, ds)
        return stc, act

    p = lam.Pipeline(SHAPE, lam.MODE_ADAPTIVE)
    p.set_weights(w16)
    p.set_fixed_k(np.full((C, 2), K, np.int32))
    p.learn_init(KW, 256, seed=1)
    p.learn_capture(NTR)
    for b in range(NTR):
        gen(b)
        p.execute(dr, ds, on_device=True)
        if (b + 1) % 4 == 0:
            p.dump_visibilities_i64()
    p.sync()
    assert p.learn_captured() == NTR
    loss = p.learn_train(8, batch=512, lr=1e-3, seed=3)
    _log(f"cfg5: training loss {loss[0]:.4f} -> {loss[-1]:.4f}")
    assert loss[-1] < loss[0]

    evecs = np.zeros((C, 2, R, R, 2), np.float32)
    counts = np.zeros((C, 2), np.int32)
    res = {}
    for name, method in (("plain", lam.EIG_PLAIN), ("previous", lam.EIG_PREVIOUS),
                         ("learned", lam.EIG_LEARNED)):
        p.set_eig_method(lam.EIG_PLAIN)
        gen(NTR - 1)
        p.execute(dr, ds, on_device=True)  # seeds the warm-start chain
        p.sync()
        p.dump_visibilities_i64()
        p.set_eig_method(method, refine_sweeps=2, residual_tol=0.1)
        p.set_profiling(True)
        p.kernel_times()
        depth, fb = [], []
        for b in range(NTR, NTR + NEV):
            stc, act = gen(b)
            p.execute(dr, ds, evecs_out=evecs, counts_out=counts, on_device=True)
            p.sync()
            _, fbb, orth = p.eig_stats(with_orthogonality=True)
            fb.append(fbb.mean())
            assert orth.max() < 1e-4, (name, b, float(orth.max()))
            U = evecs[..., 0] + 1j * evecs[..., 1]
            for c in range(C):
                a = stc[c][act[c]]
                for pp in range(2):
                    depth.append(syn.null_depth_db(W[c, pp], U[c, pp].T, counts[c, pp], a)
                                 if len(a) else np.nan)
            p.dump_visibilities_i64()
        kt = p.kernel_times()
        p.set_profiling(False)
        stage = sum(kt[k][0] for k in ("eig_null", "eig_fallback", "features", "mlp") if k in kt) / NEV
        res[name] = (np.array(depth), float(np.mean(fb)), stage)
        _log(f"cfg5: {name}: eig stage {stage * 1e3:.1f} us, fallback {res[name][1]:.4f}")
    p.close()
    dp = res["plain"][0]
    ok = ~np.isnan(dp)
    for name in ("previous", "learned"):
        diff = res[name][0][ok] - dp[ok]
        _log(f"cfg5: {name}: null depth diff mean {diff.mean():+.4f} dB, |diff| p99 "
             f"{np.quantile(np.abs(diff), 0.99):.3f} dB, stage {res[name][2] * 1e3:.1f} us vs plain "
             f"{res['plain'][2] * 1e3:.1f} us")
        assert res[name][1] <= 0.02, name
        assert abs(diff.mean()) <= 0.5, name
        assert res[name][2] < res["plain"][2], name
