"""The allocation-free correlator (lam_correlate_noalloc, static device
scratch) against the allocating entry point lam_correlate, through the C ABI:
bit-exact integer visibilities on two calls in a row, the second with a
smaller block, plus a block whose automatic split choice exceeds the scratch
(fewer splits are used) and one that cannot fit at all (rejected)."""
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


def _vis_shape(s):
    return (s.C, s.R * (s.R + 1) // 2, 2, 2, 2)


def _both(shape, window, seed):
    rng = np.random.default_rng(seed)
    ring = rng.integers(-128, 128, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    scales = np.ones(shape.scale_shape(), np.int16)
    d_ring, d_scales = lam.DeviceArray.from_numpy(ring), lam.DeviceArray.from_numpy(scales)
    zeros = np.zeros(_vis_shape(shape), np.int64)
    d_ref, d_new = lam.DeviceArray.from_numpy(zeros), lam.DeviceArray.from_numpy(zeros)
    d_flags = lam.DeviceArray.from_numpy(np.zeros(1, np.int32))
    lam.correlate(shape, window, d_ring, d_scales, d_ref)
    lam.correlate_noalloc(shape, window, d_ring, d_scales, d_new, d_flags)
    new = d_new.numpy()  # synchronous copy: ordered after the enqueued kernels
    return ring, scales, d_ref.numpy(), new, int(d_flags.numpy()[0])


def test_noalloc_matches_allocating_entry_two_calls():
    big = syn.BlockShape(C=16, F=1, NT=64, R=36, P=2, RPP=36, NP=32, B=1)
    small = syn.BlockShape(C=2, F=3, NT=16, R=12, P=2, RPP=4, NP=3, B=1)
    for i, (shape, window) in enumerate([(big, 0), (small, 1)]):
        _log(f"noalloc call {i}: {shape} window {window}")
        ring, scales, ref, new, flag = _both(shape, window, 100 + i)
        assert np.array_equal(new, ref)
        assert flag == 0
        s = orc.Shape(C=shape.C, F=shape.F, NT=shape.NT, R=shape.R, P=shape.P, RPP=shape.RPP, NP=shape.NP,
                      B=shape.B)
        w = orc.WINDOW_ADAPTIVE if window else orc.WINDOW_CORR_BEAM
        assert np.array_equal(new, orc.correlate(s, orc.ingest(s, ring, scales, w)))


def test_noalloc_fewer_splits_when_scratch_is_short():
    # the automatic choice (4 splits x 13.3 MB) exceeds the 32 MiB scratch: 2 splits are used
    shape = syn.BlockShape(C=200, F=1, NT=64, R=64, P=2, RPP=64, NP=16, B=1)
    _log(f"noalloc reduced splits: {shape}")
    _, _, ref, new, _ = _both(shape, 0, 7)
    assert np.array_equal(new, ref)


def test_noalloc_rejects_blocks_beyond_scratch():
    shape = syn.BlockShape(C=505, F=1, NT=8, R=64, P=2, RPP=64, NP=1, B=1)
    _log(f"noalloc rejection: {shape}")
    d_ring = lam.DeviceArray(shape.ring_shape(), np.int8)
    d_scales = lam.DeviceArray(shape.scale_shape(), np.int16)
    d_vis = lam.DeviceArray(_vis_shape(shape), np.int64)
    with pytest.raises(lam.LamError) as ei:
        lam.correlate_noalloc(shape, 0, d_ring, d_scales, d_vis)
    assert ei.value.code == lam.LAM_ERR_UNSUPPORTED
