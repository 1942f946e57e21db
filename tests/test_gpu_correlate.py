"""GPU parity of the kernel-level correlator entry point (lam_correlate)
against the CPU oracle: exact integer visibilities, added into the caller's
int64 buffer, for both ingest windows (corr-beam ring slots 0..NP-1,
spatial.cuh:254-264; adaptive window from slot 1, spatial.cuh:1431-1436)."""
import sys

import numpy as np
import pytest

import lam
import lam_synth as syn
import oracle as orc

pytestmark = pytest.mark.gpu


def _o(s):
    return orc.Shape(C=s.C, F=s.F, NT=s.NT, R=s.R, P=s.P, RPP=s.RPP, NP=s.NP, B=s.B)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _vis_shape(s):
    return (s.C, s.R * (s.R + 1) // 2, 2, 2, 2)


def _random_block(shape, seed):
    rng = np.random.default_rng(seed)
    ring = rng.integers(-128, 128, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    ring.reshape(-1)[::89] = -128
    scales = np.ones(shape.scale_shape(), np.int16)
    return ring, scales


@pytest.mark.parametrize("shape", [
    syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=4, B=1),
    syn.BlockShape(C=3, F=4, NT=16, R=36, P=2, RPP=9, NP=5, B=1),
    syn.BlockShape(C=2, F=1, NT=64, R=64, P=2, RPP=64, NP=3, B=1),
], ids=lambda s: f"C{s.C}F{s.F}NT{s.NT}R{s.R}NP{s.NP}")
@pytest.mark.parametrize("window", [0, 1])
def test_correlate_entry_matches_oracle(shape, window):
    _log(f"lam_correlate {shape} window {window}")
    ring, scales = _random_block(shape, 7 + shape.R + window)
    ring2, _ = _random_block(shape, 91 + shape.C)
    d_ring, d_ring2 = lam.DeviceArray.from_numpy(ring), lam.DeviceArray.from_numpy(ring2)
    d_scales = lam.DeviceArray.from_numpy(scales)
    d_vis = lam.DeviceArray.from_numpy(np.zeros(_vis_shape(shape), np.int64))
    d_flags = lam.DeviceArray.from_numpy(np.zeros(1, np.int32))
    lam.correlate(shape, window, d_ring, d_scales, d_vis, d_flags)
    s = _o(shape)
    w = orc.WINDOW_ADAPTIVE if window else orc.WINDOW_CORR_BEAM
    ref1 = orc.correlate(s, orc.ingest(s, ring, scales, w))
    assert np.array_equal(d_vis.numpy(), ref1)
    # a second block adds into the same buffer
    lam.correlate(shape, window, d_ring2, d_scales, d_vis, d_flags)
    ref2 = orc.correlate(s, orc.ingest(s, ring2, scales, w))
    assert np.array_equal(d_vis.numpy(), ref1 + ref2)
    assert d_flags.numpy()[0] == 0


def test_correlate_entry_longest_block_stays_exact():
    """65536 samples of -128 everywhere: an autocorrelation sums
    2 * 128^2 * 65536 = 2^31, one more than int32 holds; the time splits keep
    every int32 partial below 2^31 and the int64 result is exact."""
    shape = syn.BlockShape(C=1, F=1, NT=64, R=4, P=2, RPP=4, NP=1024, B=1)
    _log(f"lam_correlate longest block {shape}")
    ring = np.full(shape.ring_shape(), -128, np.int8)
    scales = np.ones(shape.scale_shape(), np.int16)
    d_ring, d_scales = lam.DeviceArray.from_numpy(ring), lam.DeviceArray.from_numpy(scales)
    d_vis = lam.DeviceArray.from_numpy(np.zeros(_vis_shape(shape), np.int64))
    lam.correlate(shape, 0, d_ring, d_scales, d_vis)
    vis = d_vis.numpy()
    s = _o(shape)
    assert np.array_equal(vis, orc.correlate(s, orc.ingest(s, ring, scales, orc.WINDOW_CORR_BEAM)))
    assert vis[0, 0, 0, 0, 0] == 2 ** 31


def test_correlate_entry_range_flag_and_errors():
    shape = syn.BlockShape(C=1, F=1, NT=8, R=4, P=2, RPP=4, NP=1, B=1, R_pad=32)
    ring, scales = syn.ring_from_fn(shape, lambda *a: (100, 0), lambda *a: 3)
    d_ring, d_scales = lam.DeviceArray.from_numpy(ring), lam.DeviceArray.from_numpy(scales)
    d_vis = lam.DeviceArray.from_numpy(np.zeros(_vis_shape(shape), np.int64))
    d_flags = lam.DeviceArray.from_numpy(np.zeros(1, np.int32))
    lam.correlate(shape, 0, d_ring, d_scales, d_vis, d_flags)
    assert d_flags.numpy()[0] == 1
    with pytest.raises(lam.LamError) as ei:
        lam.correlate(shape, 2, d_ring, d_scales, d_vis)
    assert ei.value.code == lam.LAM_ERR_ARG
    bad = syn.BlockShape(C=1, F=1, NT=8, R=65, P=2, RPP=65, NP=1, B=1)
    with pytest.raises(lam.LamError) as ei:
        lam.correlate(bad, 0, d_ring, d_scales, d_vis)
    assert ei.value.code == lam.LAM_ERR_UNSUPPORTED
