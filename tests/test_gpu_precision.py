"""Precision sweep of the beamformer's weight operands (FP16 hi+lo, FP16,
BF16) on the adaptive pipeline: 3 interferers per channel at INR 0-30 dB,
rank-3 null, against the FP64 oracle (beams) and LAPACK (eigenvalues).

Bars: hi+lo beams within 1e-3 normwise relative error (the path's stated
tolerance); FP16 and BF16 are reported with looser bounds (5e-3, 3e-2) and
must order hi+lo <= FP16 <= BF16; eigenvalues of the nulled subspace within
1e-5 relative for every mode (the eigen stage does not depend on the weight
precision)."""
import numpy as np
import pytest

import lam
import lam_synth as syn
import oracle as orc

pytestmark = pytest.mark.gpu

MODES = [("f16_hilo", lam.WPREC_F16_HILO, 1e-3), ("f16", lam.WPREC_F16, 5e-3),
         ("bf16", lam.WPREC_BF16, 3e-2)]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if lam.device_count() < 1:
        pytest.skip("no gfx950 device")


def _scene_block(shape, seed):
    sc = syn.DriftScenario(shape.C, shape.R, seed=seed, n_blocks=1, n_src=(3, 3), inr_db=(0.0, 30.0),
                           drift=0.0, p_switch=0.0, max_sources=3)
    tones, st, gain, stc, act = sc.block(0)
    dr = lam.DeviceArray(shape.ring_shape(), np.int8)
    ds = lam.DeviceArray(shape.scale_shape(), np.int16)
    dt, dst, dg = (lam.DeviceArray.from_numpy(a) for a in (tones, st, gain))
    lam.synth_block(shape, seed, 0, 16.0, dt, dst, dg, 3, dr, ds)
    return dr.numpy(), ds.numpy()


def test_weight_precision_sweep():
    shape = syn.BlockShape(C=6, F=1, NT=64, R=36, P=2, RPP=36, NP=16, B=16)
    ring, scales = _scene_block(shape, seed=31)
    w16, w = syn.random_weights(shape, seed=32)
    s = orc.Shape(C=shape.C, F=1, NT=64, R=36, P=2, RPP=36, NP=16, B=16)
    y = orc.ingest(s, ring, scales, orc.WINDOW_ADAPTIVE)
    ref = orc.adaptive_block(s, y, w, fixed_k=np.full((shape.C, 2), 3))
    den = np.max(np.abs(ref["beams2b"]))
    errs = {}
    for name, mode, bound in MODES:
        p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
        p.set_weights(w16)
        p.set_fixed_k(np.full((shape.C, 2), 3, np.int32))
        p.set_weight_precision(mode)
        out = p.run_host(ring, scales)
        p.close()
        eb = float(np.max(np.abs(out["beams"] - ref["beams2b"])) / den)
        ev = out["evals_out"]
        ee = float(np.max(np.abs(ev[..., -3:] - ref["evals"][..., -3:]) / np.abs(ref["evals"][..., -3:])))
        errs[name] = (eb, ee)
        print(f"{name:9s} beam rel err {eb:.2e}  top-3 eigenvalue rel err {ee:.2e}", flush=True)
        assert eb < bound, name
        assert ee < 1e-5, name
    assert errs["f16_hilo"][0] <= errs["f16"][0] * 1.01 + 1e-6
    assert errs["f16"][0] <= errs["bf16"][0] * 1.01 + 1e-6


def test_corr_beam_bf16_weights():
    """Corr-beam with BF16 operands: visibilities stay exact, beams within the
    BF16 bound; FP16 mode is exact in the weights (they arrive as FP16)."""
    shape = syn.BlockShape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=4, B=8)
    rng = np.random.default_rng(9)
    ring = rng.integers(-128, 128, shape.ring_shape(), dtype=np.int16).astype(np.int8)
    scales = np.ones(shape.scale_shape(), np.int16)
    w16, w = syn.random_weights(shape, seed=10)
    s = orc.Shape(C=4, F=1, NT=64, R=36, P=2, RPP=36, NP=4, B=8)
    y = orc.ingest(s, ring, scales, orc.WINDOW_CORR_BEAM)
    ref_b, ref_v = orc.beamform(s, y, w), orc.correlate(s, y)
    den = np.max(np.abs(ref_b))
    for name, mode, bound in MODES:
        p = lam.Pipeline(shape, lam.MODE_CORR_BEAM)
        p.set_weights(w16)
        p.set_weight_precision(mode)
        out = p.run_host(ring, scales)
        vis = p.dump_visibilities_i64()
        p.close()
        assert np.array_equal(vis, ref_v)
        eb = float(np.max(np.abs(out["beams"] - ref_b)) / den)
        print(f"corr-beam {name:9s} beam rel err {eb:.2e}", flush=True)
        assert eb < bound
