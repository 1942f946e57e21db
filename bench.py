#!/usr/bin/env python3
"""Benchmark of the LAMBDA per-block correlate + beamform + RFI-null path on
one MI355X.

Workload (default): 384 channels x 2 pols x 36 antennas (64-receiver LAMBDA
layout), 8192 int8 complex samples per block, 16 beams of FP16 random-phase
weights, one CW tone per channel, plain batched Jacobi on all 768 36x36
covariances, rank-1 null, int64 visibilities accumulated every block; the
adaptive pipeline outputs 2 x 16 beams (original + nulled).

A step = one block through the adaptive pipeline with the voltages resident
in HBM.  Blocks run back to back on 3 pipeline buffers (the reference
benchmark's default, apps/bench_gpu.cu:275-279), so consecutive blocks
overlap; per-kernel durations for the roofline come from a second pass of
the same steps run one at a time.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cuda-spatial-filtering_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import lam  # noqa: E402
import lam_synth as syn  # noqa: E402

CHANNEL_RATE_HZ = 781250.0
REF_RATE_HZ = 15500.0 * 64.0  # scripts/benchmark/README.md:48-50
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_inputs(shape, seed, n_tones=1):
    scene = syn.make_scene(shape, seed=seed, n_tones=n_tones, amp=40.0)
    ring, scales = syn.make_block(shape, scene, seed=seed + 1)
    w16, w = syn.random_weights(shape, seed=seed + 2)
    return scene, ring, scales, w16, w


def algorithmic_bytes(shape, nbeams_out):
    """Bytes a kernel must move at minimum (no padding, no re-reads)."""
    C, R, T = shape.C, shape.R, shape.T
    volt = C * R * 2 * T * 2                       # int8 complex voltages, both pols
    vis = C * (R * (R + 1) // 2) * 8 * 4           # int32 visibilities [C][NBL][2][2][2]
    beams = C * 2 * nbeams_out * T * 4             # fp16 complex beams
    cov = C * 2 * R * R * 8                        # fp32 complex covariance / basis
    return dict(volt=volt, vis=vis, beams=beams, cov=cov)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(shape, ring, scales, w, channels, min_seconds=10.0, max_passes=4):
    """FP64 oracle on the same block, `channels` channels at a time over the
    whole band, repeated until about `min_seconds` of host work (all host
    cores through the oracle's OpenMP loops), then one call of `channels`
    channels on one thread (the single-core figure)."""
    import oracle as orc

    s = orc.Shape(C=channels, F=shape.F, NT=shape.NT, R=shape.R, P=2, RPP=shape.RPP,
                  NP=shape.NP, B=shape.B)
    threads = orc.num_threads()

    def one(c0):
        y = orc.ingest(s, ring[c0:c0 + channels], scales[c0:c0 + channels], orc.WINDOW_ADAPTIVE)
        orc.adaptive_block(s, y, w[c0:c0 + channels], fixed_k=np.ones((channels, 2)))

    done = 0
    t0 = time.perf_counter()
    passes = 0
    while passes < max_passes and (passes == 0 or time.perf_counter() - t0 < min_seconds):
        for c0 in range(0, shape.C - channels + 1, channels):
            one(c0)
            done += channels
        passes += 1
    dt = time.perf_counter() - t0
    orc.set_threads(1)
    t1 = time.perf_counter()
    one(0)
    dt1 = time.perf_counter() - t1
    orc.set_threads(threads)
    per_channel = shape.R * shape.T
    model = cpu_model()
    return dict(value=per_channel * done / dt, unit="ant-samples/s", cores=threads,
                kind="port",
                sample=f"{passes} pass(es) over the {shape.C}-channel block of {shape.T} samples "
                       f"({done} channel-blocks, {channels} channels per call), FP64 oracle: "
                       f"ingest+correlate+whiten+Jacobi+null+W.P+{2 * shape.B} beams; "
                       f"{dt:.2f} s wall on {threads} threads of {model}",
                seconds=dt, cpu_model=model,
                single_core={"value": per_channel * channels / dt1, "unit": "ant-samples/s",
                             "cores": 1, "seconds": dt1,
                             "sample": f"one call of {channels} channels of the same block on one thread"})


def run_spectra(p, shape, d_ring, d_scales, args):
    """Adaptive pipeline with the post-beam spectra on (the reference's
    register_fft_block product): streaming ms/step, and the spectra kernel's
    own duration against the HBM bytes it must move (read the 2B fp16 beams,
    write the [C][2][2B][54] float spectra)."""
    n = max(10, args.steps // 3)
    p.set_spectra(True)
    for _ in range(4):
        p.execute(d_ring, d_scales, on_device=True)
    p.sync()
    p.dump_visibilities_i64()
    t0 = time.perf_counter()
    for i in range(n):
        p.execute(d_ring, d_scales, on_device=True)
    p.sync()
    t1 = time.perf_counter()
    p.dump_visibilities_i64()
    p.set_profiling(True)
    p.kernel_times()
    for i in range(n):
        p.execute(d_ring, d_scales, on_device=True)
        p.sync()
    kt = p.kernel_times()
    p.set_profiling(False)
    p.set_spectra(False)
    p.dump_visibilities_i64()
    ms, launches = kt.get("spectra", (0.0, 0))
    k_ms = ms / max(1, launches)
    nbytes = shape.C * 2 * 2 * shape.B * shape.T * 4 + shape.C * 2 * 2 * shape.B * (shape.NT - 10) * 4
    gbs = nbytes / (k_ms * 1e-3) / 1e9 if k_ms > 0 else None
    sec = (t1 - t0) / n
    return {"ms_per_step": sec * 1e3, "real_time_factor": (shape.T / CHANNEL_RATE_HZ) / sec,
            "kernel_ms": k_ms, "algorithmic_bytes": nbytes,
            "GBps_algorithmic": None if gbs is None else round(gbs, 1),
            "hbm_frac": None if gbs is None else round(gbs / HBM_PEAK_GBS, 4)}


def run_host_fed(p, shape, ring, scales, args):
    """Blocks from pinned host buffers recycled on the pipeline's release
    callback (the reference's ProcessorState -> execute_pipeline path): each
    step includes the block's H2D copy, overlapped with the previous blocks'
    kernels on the pipeline's buffers.  The buffers keep the same synthetic
    block (a reassembly engine writes them in place as packets arrive, so its
    copy is not part of a step)."""
    n = max(8, args.steps // 4)
    hr = lam.HostRing(p, args.buffers)
    for i in range(args.buffers):
        hr.samples[i].array[...] = ring
        hr.scales[i].array[...] = scales
    for _ in range(args.buffers):
        hr.submit(hr.acquire())
    p.sync()
    p.dump_visibilities_i64()
    t0 = time.perf_counter()
    for _ in range(n):
        hr.submit(hr.acquire())
    p.sync()
    t1 = time.perf_counter()
    p.dump_visibilities_i64()
    hr.close()
    sec = (t1 - t0) / n
    nbytes = int(np.prod(shape.ring_shape())) + 2 * int(np.prod(shape.scale_shape()))
    return {"ms_per_step": sec * 1e3, "real_time_factor": (shape.T / CHANNEL_RATE_HZ) / sec,
            "h2d_bytes_per_step": nbytes, "h2d_GBps_effective": round(nbytes / sec / 1e9, 1),
            "steps": n, "buffers": args.buffers}


def run_lambda_gpu(shape, d_ring, d_scales, w16, args):
    """LambdaGPUPipeline on the same block: B beams, accumulated visibilities,
    the eigendecomposition of all four pol-pair covariances of every block
    (4 x 384 = 1536 problems of 36 x 36) and the bandpass (8192-point FFT of
    every beam row, |X|^2 averaged over runs of 128 bins)."""
    n = max(10, args.steps // 3)
    p = lam.Pipeline(shape, lam.MODE_LAMBDA_GPU, num_buffers=args.buffers)
    p.set_weights(w16)
    p.set_eig_params(args.sweeps, 1e-6)
    p.set_bandpass(128)  # FFT_DOWNSAMPLE_FACTOR default of LambdaConfig
    p.set_tuning(args.spw)
    for _ in range(5):
        p.execute(d_ring, d_scales, on_device=True)
    p.sync()
    p.dump_visibilities_i64()
    t0 = time.perf_counter()
    for _ in range(n):
        p.execute(d_ring, d_scales, on_device=True)
    p.sync()
    t1 = time.perf_counter()
    p.dump_visibilities_i64()
    p.set_profiling(True)
    p.kernel_times()
    for _ in range(n):
        p.execute(d_ring, d_scales, on_device=True)
        p.sync()
    kt = p.kernel_times()
    p.close()
    sec = (t1 - t0) / n
    return {"ms_per_step": sec * 1e3, "real_time_factor": (shape.T / CHANNEL_RATE_HZ) / sec,
            "ant_samples_per_s": shape.R * shape.C * shape.T / sec,
            "kernels_us": {k: round(v[0] / max(1, v[1]) * 1e3, 1) for k, v in kt.items()},
            "eigenproblems_per_block": shape.C * 4}


def run_eigen_outputs(shape, d_ring, d_scales, w16, args):
    """The headline workload with every block's eigenvalues and eigenvectors
    landing in device buffers: such blocks take the cyclic Jacobi solver
    (all n eigenvectors), where blocks without eigen outputs take the
    tridiagonal solve for the removed modes only."""
    n = max(10, args.steps // 3)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE, num_buffers=args.buffers)
    p.set_weights(w16)
    p.set_fixed_k(np.ones((shape.C, 2), np.int32))
    p.set_eig_params(args.sweeps, 1e-6)
    p.set_tuning(args.spw)
    evals = lam.DeviceArray((shape.C, 2, shape.R), np.float32)
    evecs = lam.DeviceArray((shape.C, 2, shape.R, shape.R, 2), np.float32)
    for _ in range(5):
        p.execute(d_ring, d_scales, evals_out=evals, evecs_out=evecs, on_device=True)
    p.sync()
    p.dump_visibilities_i64()
    t0 = time.perf_counter()
    for _ in range(n):
        p.execute(d_ring, d_scales, evals_out=evals, evecs_out=evecs, on_device=True)
    p.sync()
    t1 = time.perf_counter()
    p.dump_visibilities_i64()
    p.set_profiling(True)
    p.kernel_times()
    for _ in range(n):
        p.execute(d_ring, d_scales, evals_out=evals, evecs_out=evecs, on_device=True)
        p.sync()
    kt = p.kernel_times()
    sweeps = float(p.eig_stats()[0].mean())
    p.close()
    sec = (t1 - t0) / n
    return {"ms_per_step": sec * 1e3, "real_time_factor": (shape.T / CHANNEL_RATE_HZ) / sec,
            "ant_samples_per_s": shape.R * shape.C * shape.T / sec, "solver": "cyclic Jacobi",
            "sweeps_mean": round(sweeps, 2),
            "kernels_us": {k: round(v[0] / max(1, v[1]) * 1e3, 1) for k, v in kt.items()}}


def run_projection(shape, d_ring, d_scales, args, k=1, runs=4):
    """LambdaProjectionPipeline on the same block: per block the projector onto
    the K largest eigenvectors of all four pol-pair covariances is accumulated
    (1536 problems), and every `runs` blocks the average is decomposed (the
    dump is inside the timed loop at that cadence)."""
    n = max(12, (args.steps // 3) // runs * runs)
    p = lam.Pipeline(shape, lam.MODE_PROJECTION, num_buffers=args.buffers)
    p.set_projection(k, runs)
    p.set_eig_params(args.sweeps, 1e-6)
    p.set_tuning(args.spw)
    for _ in range(runs):
        p.execute(d_ring, d_scales, on_device=True)
    p.dump_projection()
    ev = lam.HostArray((shape.C, 2, 2, shape.R), np.float32)
    vec = lam.HostArray((shape.C, 2, 2, shape.R, shape.R, 2), np.float32)
    t0 = time.perf_counter()
    for i in range(n):
        p.execute(d_ring, d_scales, on_device=True)
        if (i + 1) % runs == 0:
            p.dump_projection(wait=False, out=(ev.array, vec.array))
    p.sync()
    t1 = time.perf_counter()
    p.set_profiling(True)
    p.kernel_times()
    for _ in range(runs):
        p.execute(d_ring, d_scales, on_device=True)
        p.sync()
    kt = p.kernel_times()
    p.dump_projection()
    p.close()
    ev.free()
    vec.free()
    sec = (t1 - t0) / n
    return {"ms_per_step": sec * 1e3, "real_time_factor": (shape.T / CHANNEL_RATE_HZ) / sec,
            "nr_signal_eigenvectors": k, "nr_runs_to_average": runs,
            "kernels_us": {k2: round(v[0] / max(1, v[1]) * 1e3, 1) for k2, v in kt.items()}}


def run_learned(args):
    """Learned warm start on 384 ch x 2 pol x 36 ant with drifting, switching
    interferers: blocks [0, train) run the plain solver and are captured as the
    training set; the MLP is trained on the GPU; blocks [train, total) run with
    the plain solver, the previous block's vectors as warm start, and the MLP's
    proposal, on the same seeded voltages.  Reports the eigen-stage time
    (features + MLP + warm solve + deferred fallback, or the plain solve), the
    fallback rate and the null depth against the known steering vectors."""
    C, KW, K = args.channels, 8, 4
    shape = syn.BlockShape(C=C, F=1, NT=64, R=36, P=2, RPP=36, NP=128, B=16, blocks_to_accumulate=4)
    nb, ntr = args.learned_blocks, args.train_blocks
    sc = syn.DriftScenario(C, 36, seed=20261203 + 4, n_blocks=nb)
    w16, W = syn.random_weights(shape, seed=20261203 + 5)
    dr = lam.DeviceArray(shape.ring_shape(), np.int8)
    ds = lam.DeviceArray(shape.scale_shape(), np.int16)
    dt = lam.DeviceArray((C, 4, 6), np.float32)
    dst = lam.DeviceArray((C, 4, 36, 2), np.float32)
    dg = lam.DeviceArray((C,), np.float32)

    def gen(b, ring=dr, scales=ds):
        tones, st, gain, stc, act = sc.block(b)
        dt.upload(tones)
        dst.upload(st)
        dg.upload(gain)
        lam.synth_block(shape, 20261203 + 4, b, 16.0, dt, dst, dg, 4, ring, scales)
        return stc, act

    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE, num_buffers=args.buffers)
    p.set_weights(w16)
    p.set_fixed_k(np.full((C, 2), K, np.int32))
    p.set_eig_params(args.sweeps, 1e-6)
    p.set_graphs(False)
    p.learn_init(KW, 256, seed=1)
    p.learn_capture(ntr)
    log(f"learned: capturing {ntr} plain blocks ...")
    for b in range(ntr):
        gen(b)
        p.execute(dr, ds, on_device=True)
        if (b + 1) % 4 == 0:
            p.dump_visibilities_i64()
    p.sync()
    log(f"learned: training {args.epochs} epochs on {p.learn_captured()} blocks ...")
    t0 = time.perf_counter()
    loss = p.learn_train(args.epochs, batch=512, lr=1e-3, seed=7)
    train_s = time.perf_counter() - t0
    evecs = np.zeros((C, 2, 36, 36, 2), np.float32)
    counts = np.zeros((C, 2), np.int32)
    stage_keys = ("eig_null", "eig_fallback", "features", "mlp")
    # autotuner inputs: the last training-period blocks (never the evaluated ones)
    tune = []
    if not args.no_autotune:
        for b in range(max(0, ntr - 8), ntr):
            gen(b)
            r2 = lam.DeviceArray(shape.ring_shape(), np.int8)
            s2 = lam.DeviceArray(shape.scale_shape(), np.int16)
            lam.memcpy(r2, dr, dr.nbytes)
            lam.memcpy(s2, ds, ds.nbytes)
            tune.append((r2, s2))
    results = {}
    evalset = None
    for name, method in (("plain", lam.EIG_PLAIN), ("previous", lam.EIG_PREVIOUS),
                         ("learned", lam.EIG_LEARNED)):
        steps, tuned = args.refine, None
        if tune:
            p.set_eig_method(method, refine_sweeps=args.refine, residual_tol=0.1)
            tuned = p.autotune([t[0] for t in tune], [t[1] for t in tune], iters=2)
            steps = tuned["refine_steps"]
            log(f"learned: {name}: autotune {tuned}")
        p.set_eig_method(lam.EIG_PLAIN)
        gen(ntr - 1)
        p.execute(dr, ds, on_device=True)  # seeds the warm-start chain
        p.sync()
        p.set_eig_method(method, refine_sweeps=steps, residual_tol=0.1)
        depth, fb, orth = [], [], []
        for b in range(ntr, nb):
            stc, act = gen(b)
            p.execute(dr, ds, evecs_out=evecs, counts_out=counts, on_device=True)
            p.sync()
            _, fbb, orb = p.eig_stats(with_orthogonality=True)
            orth.append(float(orb.max()))
            if method != lam.EIG_PLAIN:
                fb.append(fbb.mean())
            U = evecs[..., 0] + 1j * evecs[..., 1]
            for c in range(C):
                a = stc[c][act[c]]
                for pp in range(2):
                    depth.append(syn.null_depth_db(W[c, pp], U[c, pp].T, counts[c, pp], a)
                                 if len(a) else np.nan)
        # the evaluated blocks again, resident in HBM, for timing
        if evalset is None:
            evalset = []
            for b in range(ntr, nb):
                r2 = lam.DeviceArray(shape.ring_shape(), np.int8)
                s2 = lam.DeviceArray(shape.scale_shape(), np.int16)
                gen(b, r2, s2)
                evalset.append((r2, s2))

        def seed_chain():
            p.set_eig_method(lam.EIG_PLAIN)
            gen(ntr - 1)
            p.execute(dr, ds, on_device=True)
            p.sync()
            p.dump_visibilities_i64()
            p.set_eig_method(method, refine_sweeps=steps, residual_tol=0.1)

        # per-kernel durations: one block at a time, HIP events around each launch
        seed_chain()
        p.set_profiling(True)
        p.kernel_times()
        for r2, s2 in evalset:
            p.execute(r2, s2, on_device=True)
            p.sync()
        kt = p.kernel_times()
        p.set_profiling(False)
        # streaming throughput: blocks back to back, buffers overlap, graphs replay each phase
        seed_chain()
        p.set_graphs(True)
        t0 = time.perf_counter()
        for r2, s2 in evalset:
            p.execute(r2, s2, on_device=True)
        p.sync()
        stream_ms = (time.perf_counter() - t0) / len(evalset) * 1e3
        fb_stream = float(p.eig_stats()[1].mean())
        p.set_graphs(False)
        p.dump_visibilities_i64()
        n_eval = nb - ntr
        per = {k: v[0] / v[1] * 1e3 for k, v in kt.items()}
        stage_us = sum(kt[k][0] for k in stage_keys if k in kt) / n_eval * 1e3
        gpu_ms = sum(v[0] for v in kt.values()) / n_eval
        results[name] = dict(eig_stage_us=round(stage_us, 1), gpu_ms_per_block=round(gpu_ms, 4),
                             streaming_ms_per_block=round(stream_ms, 4),
                             real_time_factor=round((shape.T / CHANNEL_RATE_HZ) / (stream_ms * 1e-3), 3),
                             ant_samples_per_s=shape.R * C * shape.T / (stream_ms * 1e-3),
                             streaming_last_block_fallback_rate=fb_stream,
                             autotune=tuned,
                             kernels_us={k: round(v, 1) for k, v in per.items()},
                             fallback_rate=float(np.mean(fb)) if fb else 0.0,
                             null_basis_orthogonality_fp64_max=max(orth),
                             depth=np.array(depth))
        log(f"learned: {name}: eig stage {stage_us:.1f} us, streaming {stream_ms:.3f} ms/block")
    p.close()
    for r2, s2 in tune + (evalset or []):
        r2.free()
        s2.free()
    plain_depth = results["plain"]["depth"]
    ok = ~np.isnan(plain_depth)
    out = dict(workload=f"{C} ch x 2 pol x 36 ant, {nb} blocks of {shape.T} samples, 1-4 drifting "
                        f"CW interferers (INR 5-30 dB, 0.01 rad/block, on/off p=0.1), rank-{K} null; "
                        f"train blocks 0-{ntr - 1}, evaluate {ntr}-{nb - 1}",
               warm_vectors=KW, epochs=args.epochs,
               train_seconds=round(train_s, 2), loss_first_last=[float(loss[0]), float(loss[-1])])
    for name, r in results.items():
        d = r.pop("depth")
        r["null_depth_db_mean"] = round(float(np.mean(d[ok])), 3)
        if name != "plain":
            diff = d[ok] - plain_depth[ok]
            r["null_depth_diff_db_mean"] = round(float(np.mean(diff)), 4)
            r["null_depth_diff_db_abs_p99"] = round(float(np.quantile(np.abs(diff), 0.99)), 4)
            r["speedup_vs_plain"] = round(results["plain"]["eig_stage_us"] / r["eig_stage_us"], 3)
        out[name] = r
    return out


def run_precision(args):
    """Beam-weight precision sweep on the cfg2 shape: 3 CW interferers per
    channel at INR 0-30 dB, rank-3 null.  Per precision: the 2B-beam kernel
    time, the end-to-end null depth measured on the output beams (each known
    tone's amplitude in the nulled beams against the original beams), the
    beams' difference from the hi+lo run, and the top-3 eigenvalues against
    LAPACK on the block's own exact visibilities."""
    C, K = args.channels, 3
    shape = syn.BlockShape(C=C, F=1, NT=64, R=36, P=2, RPP=36, NP=128, B=16, blocks_to_accumulate=1)
    seed = 20261203 + 3
    sc = syn.DriftScenario(C, 36, seed=seed, n_blocks=1, n_src=(3, 3), inr_db=(0.0, 30.0), drift=0.0,
                           p_switch=0.0, max_sources=3)
    tones, st, gain, _, _ = sc.block(0)
    dr = lam.DeviceArray(shape.ring_shape(), np.int8)
    ds = lam.DeviceArray(shape.scale_shape(), np.int16)
    dt, dst, dg = (lam.DeviceArray.from_numpy(a) for a in (tones, st, gain))
    lam.synth_block(shape, seed, 0, 16.0, dt, dst, dg, 3, dr, ds)
    w16, _ = syn.random_weights(shape, seed=seed + 1)
    T, B = shape.T, shape.B
    t = np.arange(T)
    # tone phasors e^{-2 pi i nu t} per (channel, source)
    ph = np.exp(-2j * np.pi * tones[..., 1].astype(np.float64)[..., None] * t)  # [C][S][T]
    out, ref_beams = {}, None
    for name, mode in (("f16_hilo", lam.WPREC_F16_HILO), ("f16", lam.WPREC_F16), ("bf16", lam.WPREC_BF16)):
        p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
        p.set_weights(w16)
        p.set_fixed_k(np.full((C, 2), K, np.int32))
        p.set_weight_precision(mode)
        beams = p.beams_array()
        evals = np.zeros((C, 2, 36), np.float32)
        p.execute(dr, ds, beams_out=beams, evals_out=evals, on_device=True)
        p.sync()
        vis = p.dump_visibilities_i64()
        p.set_profiling(True)
        p.kernel_times()
        for _ in range(5):
            p.execute(dr, ds, on_device=True)
        kt = p.kernel_times()
        p.close()
        y = beams[..., 0].astype(np.float64) + 1j * beams[..., 1].astype(np.float64)  # [C][2][2B][T]
        amp = np.einsum("cpbt,cst->cpbs", y, ph) / T
        orig = np.sum(np.abs(amp[:, :, :B]) ** 2, axis=(2, 3))
        mitig = np.sum(np.abs(amp[:, :, B:]) ** 2, axis=(2, 3))
        depth = 10 * np.log10(np.maximum(mitig, 1e-30) / orig)
        # LAPACK on the whitened XX / YY covariance of the block's exact visibilities
        Vc = vis[..., 0] + 1j * vis[..., 1]  # [C][nbl][2][2]
        ii, jj = np.triu_indices(36)
        bl = jj * (jj + 1) // 2 + ii
        A = np.zeros((C, 2, 36, 36), np.complex128)
        for pp in range(2):
            A[:, pp, ii, jj] = Vc[:, bl, pp, pp]
            A[:, pp, jj, ii] = np.conj(Vc[:, bl, pp, pp])
        dd = np.sqrt(np.maximum(np.einsum("cpii->cpi", A).real, 1e-10))
        lam_ref = np.linalg.eigvalsh(A / (dd[..., :, None] * dd[..., None, :]))
        ev_err = np.abs(evals[..., -K:] - lam_ref[..., -K:]) / np.abs(lam_ref[..., -K:])
        r = dict(beam_2b_us=round(kt["beam_2b"][0] / kt["beam_2b"][1] * 1e3, 1),
                 null_depth_db_median=round(float(np.median(depth)), 2),
                 null_depth_db_p90=round(float(np.quantile(depth, 0.9)), 2),
                 top3_eigenvalue_rel_err_max=float(ev_err.max()))
        if ref_beams is None:
            ref_beams = y
        else:
            r["beam_rel_diff_vs_f16_hilo"] = float(np.max(np.abs(y - ref_beams)) / np.max(np.abs(ref_beams)))
        out[name] = r
        log(f"precision {name}: {r}")
    out["workload"] = (f"{C} ch x 2 pol x 36 ant, {T} samples, 3 CW interferers per channel at INR 0-30 dB, "
                       "rank-3 null, 16 -> 32 beams; beam error against the FP64 oracle is in "
                       "tests/test_gpu_precision.py")
    return out


def run_subarrays(args):
    """Uneven batching: 384 ch x 2 pol x 2 subarray groups of 12, 24 or 36
    antennas (1536 subarray eigenproblems per block) beside the 768 full-array
    problems, 1-3 CW interferers per channel at INR 10-30 dB, 8192 samples.
    The per-block sequence (correlator, one persistent launch for all 2304
    eigen / null problems, (2 + 2) x 16 beams incl. the groups' mitigated
    beams) runs from its hipGraphs; the autotuner picks the launch shapes
    first.  Reports the graph-replayed block time and the persistent launch's
    time and problem rate."""
    C, G = args.channels, 2
    shape = syn.BlockShape(C=C, F=1, NT=64, R=36, P=2, RPP=36, NP=128, B=16, blocks_to_accumulate=4)
    seed = 20261203 + 4 + 100
    sc = syn.DriftScenario(C, 36, seed=seed, n_blocks=1, n_src=(1, 3), inr_db=(10.0, 30.0), drift=0.0,
                           p_switch=0.0, max_sources=3)
    tones, st, gain, _, act = sc.block(0)
    dr = lam.DeviceArray(shape.ring_shape(), np.int8)
    ds = lam.DeviceArray(shape.scale_shape(), np.int16)
    dt, dst, dg = (lam.DeviceArray.from_numpy(a) for a in (tones, st, gain))
    lam.synth_block(shape, seed, 0, 16.0, dt, dst, dg, 3, dr, ds)
    rng = np.random.default_rng(seed)
    mask = np.zeros((C, G, 36), np.uint8)
    sizes = rng.choice([12, 24, 36], (C, G))
    for c in range(C):
        for g in range(G):
            mask[c, g, rng.choice(36, sizes[c, g], replace=False)] = 1
    k = np.repeat(act.sum(1).astype(np.int32)[:, None, None], 2, axis=1).repeat(G, axis=2)
    p = lam.Pipeline(shape, lam.MODE_ADAPTIVE)
    p.set_weights(syn.random_weights(shape, seed=seed + 1)[0])
    p.set_fixed_k(act.sum(1).astype(np.int32)[:, None].repeat(2, axis=1))
    p.set_subarrays(mask, k)
    tuned = None if args.no_autotune else p.autotune([dr], [ds], iters=2)
    p.set_graphs(True)
    p.set_profiling(False)
    for i in range(3):
        p.execute(dr, ds, on_device=True)
    p.sync()
    p.dump_visibilities_i64()
    nblk = max(8, args.steps)
    t0 = time.perf_counter()
    for i in range(nblk):
        p.execute(dr, ds, on_device=True)
        if (i + 1) % 4 == 0:
            p.dump_visibilities_i64()
    p.sync()
    graph_ms = (time.perf_counter() - t0) / nblk * 1e3
    p.set_profiling(True)
    p.kernel_times()
    for i in range(5):
        p.execute(dr, ds, on_device=True)
    kt = p.kernel_times()
    res = p.subarray_results()
    p.close()
    key = "subarray_eig" if "subarray_eig" in kt else "eig_null"  # one launch: the queue of all problems
    sub_ms = kt[key][0] / kt[key][1]
    nprob = C * 2 * G + C * 2
    return dict(workload=f"{C} ch x 2 pol x {G} groups of 12/24/36 antennas + the full array ({nprob} "
                         f"eigenproblems in one persistent launch), {(2 + G) * 16} output beams, "
                         f"1-3 CW interferers at INR 10-30 dB, {shape.T} samples/block",
                graph_ms_per_block=round(graph_ms, 4),
                real_time_factor=round((shape.T / CHANNEL_RATE_HZ) / (graph_ms * 1e-3), 3),
                subarray_launch_us=round(sub_ms * 1e3, 1),
                eigenproblems_per_s=round(nprob / (sub_ms * 1e-3), 1),
                mean_sweeps=round(float(np.mean(res["sweeps"])), 2),
                sizes_hist={int(n): int(np.sum(res["sizes"] == n)) for n in (12, 24, 36)},
                kernels_us={kk: round(v[0] / v[1] * 1e3, 1) for kk, v in kt.items()},
                autotune=tuned)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    # untimed steps before the timed region: after the autotuner's last block the
    # overlapped stream needs ~50 blocks to reach its steady state (384 ch: 0.445 ms
    # per step timed after 10 warm-up steps, 0.432 after 50)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--channels", type=int, default=384)
    ap.add_argument("--mode", choices=["adaptive", "corr_beam"], default="adaptive")
    ap.add_argument("--sweeps", type=int, default=10)
    ap.add_argument("--cpu-channels", type=int, default=48)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graphs", type=int, default=1, help="replay each block phase from a hipGraph")
    ap.add_argument("--buffers", type=int, default=3,
                    help="pipeline buffers: consecutive blocks overlap (the reference benchmark's default is 3)")
    ap.add_argument("--learned-blocks", type=int, default=200)
    ap.add_argument("--train-blocks", type=int, default=150)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--refine", type=int, default=2, help="warm-start refinement steps")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-subarrays", action="store_true")
    ap.add_argument("--no-spectra", action="store_true")
    ap.add_argument("--no-host-fed", action="store_true")
    ap.add_argument("--no-lambda-gpu", action="store_true")
    ap.add_argument("--no-eigen-outputs", action="store_true",
                    help="skip the run with eigenvalues/eigenvectors requested (Jacobi solver)")
    ap.add_argument("--spw", type=int, default=0, help="samples per voltage workgroup (0: automatic)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run device-to-device copies of this size (HBM counter calibration)")
    args = ap.parse_args()

    if args.copy_probe_mib:
        nbytes = args.copy_probe_mib << 20
        src = lam.DeviceArray((nbytes,), np.uint8)
        dst = lam.DeviceArray((nbytes,), np.uint8)
        for _ in range(3):
            lam.memcpy(dst, src, nbytes)
        print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}), flush=True)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(channels shard with no exchange between GPUs)"}))
        sys.exit(2)

    shape = syn.BlockShape(C=args.channels, F=1, NT=64, R=36, P=2, RPP=36, NP=128, B=16,
                           blocks_to_accumulate=4)
    log(f"generating {shape.C}x{shape.T} block ...")
    scene, ring, scales, w16, w = make_inputs(shape, seed=20261203 + 2)
    d_ring = lam.DeviceArray.from_numpy(ring)
    d_scales = lam.DeviceArray.from_numpy(scales)
    mode = lam.MODE_ADAPTIVE if args.mode == "adaptive" else lam.MODE_CORR_BEAM
    p = lam.Pipeline(shape, mode, num_buffers=args.buffers)
    p.set_weights(w16)
    if mode == lam.MODE_ADAPTIVE:
        p.set_fixed_k(np.ones((shape.C, 2), np.int32))
        p.set_eig_params(args.sweeps, 1e-6)
    p.set_graphs(bool(args.graphs))
    p.set_tuning(args.spw)
    tuned = None
    if not args.no_autotune and not args.spw:
        log("autotune ...")
        tuned = p.autotune([d_ring], [d_scales], iters=3)
        log(f"autotune: {tuned}")
    p.set_profiling(False)

    log("warmup ...")
    for i in range(args.warmup):
        p.execute(d_ring, d_scales, on_device=True)
        if (i + 1) % 4 == 0:
            p.dump_visibilities_i64()
    p.sync()
    p.dump_visibilities_i64()

    # timed region: blocks back to back, buffers overlapping, no per-kernel events
    log(f"timing {args.steps} steps ...")
    p.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        p.execute(d_ring, d_scales, on_device=True)
    p.sync()
    t1 = time.perf_counter()
    p.dump_visibilities_i64()
    # per-kernel durations for the roofline: the same steps one at a time, HIP
    # events around every launch on the launching stream
    log(f"kernel timing {args.steps} serial steps ...")
    p.set_profiling(True)
    for i in range(args.steps):
        p.execute(d_ring, d_scales, on_device=True)
        p.sync()
    kt = p.kernel_times()
    p.set_profiling(False)
    orth_max = None
    sweeps_mean = None
    if mode == lam.MODE_ADAPTIVE:
        sw_last, _, orth = p.eig_stats(with_orthogonality=True)
        orth_max = float(orth.max())
        sweeps_mean = float(sw_last.mean())
    host_fed = None
    if not args.no_host_fed:
        host_fed = run_host_fed(p, shape, ring, scales, args)
        log(f"host-fed: {host_fed}")
    spectra = None
    if mode == lam.MODE_ADAPTIVE and not args.no_spectra:
        spectra = run_spectra(p, shape, d_ring, d_scales, args)
        log(f"spectra: {spectra}")
    p.close()

    sec = (t1 - t0) / args.steps
    ant_samples = shape.R * shape.C * shape.T
    value = ant_samples / sec
    rtf = (shape.T / CHANNEL_RATE_HZ) / sec
    nb_out = 2 * shape.B if mode == lam.MODE_ADAPTIVE else shape.B
    ab = algorithmic_bytes(shape, nb_out)
    per_kernel_bytes = {
        "corr": ab["volt"] + ab["vis"],
        "beam_2b": ab["volt"] + ab["beams"],
        "corr_beam": ab["volt"] + ab["beams"] + ab["vis"],
        "eig_null": ab["vis"] // 2 + ab["cov"],
    }
    kstats = {k: dict(ms_avg=v[0] / max(1, v[1]), launches=v[1]) for k, v in kt.items()}
    dominant = max(kstats, key=lambda k: kstats[k]["ms_avg"] * kstats[k]["launches"]) if kstats else None
    roofline = None
    if dominant:
        ms = kstats[dominant]["ms_avg"]
        gbs = per_kernel_bytes[dominant] / (ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dominant,
                    "kernel_ms": round(ms, 4), "algorithmic_bytes": per_kernel_bytes[dominant]}
    for k, v in kstats.items():
        v["GBps_algorithmic"] = round(per_kernel_bytes[k] / (v["ms_avg"] * 1e-3) / 1e9, 1)
    # the eigen stage is bound by neither HBM nor MFMA: report it as matrices/s and
    # as a fraction of the FP32 vector peak.  Blocks without eigen outputs take the
    # tridiagonal solve (0 sweeps): its algorithmic work is the Householder
    # reduction of a complex Hermitian matrix, 16/3 N^3 real flops (the
    # eigenvalues and the removed modes' vectors are O(N^2)); blocks with eigen
    # outputs take cyclic Jacobi: N(N-1)/2 rotations x 6N complex updates x 8
    # flops per sweep and matrix
    eig_roof = None
    if "eig_null" in kstats and sweeps_mean is not None:
        ms = kstats["eig_null"]["ms_avg"]
        nmat = shape.C * 2
        n = shape.R
        if sweeps_mean == 0:
            solver = "householder tridiagonal + multisection + twisted factorization"
            flops = nmat * 16.0 / 3.0 * n ** 3
        else:
            solver = "cyclic Jacobi"
            flops = nmat * sweeps_mean * (n * (n - 1) / 2) * 6 * n * 8
        tf = flops / (ms * 1e-3) / 1e12
        eig_roof = {"bound": "fp32 valu / latency", "solver": solver,
                    "matrices_per_s": round(nmat / (ms * 1e-3), 1),
                    "matrices": nmat, "n": n, "sweeps_mean": round(sweeps_mean, 2), "kernel_ms": round(ms, 4),
                    "achieved_tflops": round(tf, 2), "peak_tflops": 157.3, "frac": round(tf / 157.3, 4),
                    "peak_tflops_unpacked": 78.6, "frac_unpacked": round(tf / 78.6, 4)}
    # HBM bytes per launch measured with rocprofv3 PMC (FETCH_SIZE, WRITE_SIZE scaled by a copy
    # probe), per time-split count of the voltage workgroups
    tfile = os.path.join(REPO, "profiles", "adaptive_384ch_hbm_traffic.json")
    if roofline and mode == lam.MODE_ADAPTIVE and shape.C == 384 and os.path.exists(tfile):
        spw = tuned["samples_per_wg"] if tuned else args.spw
        if spw <= 0:  # the adaptive pipeline's default: about one workgroup per CU (256 CUs)
            target = max(1, -(-256 // shape.C))
            spw = -(-shape.T // target)
        spw = -(-spw // 128) * 128
        nsplit = -(-shape.T // spw)
        ent = json.load(open(tfile))["by_nsplit"].get(str(nsplit))
        if ent:
            roofline["traffic"] = ent["bytes_per_launch"].get(dominant)
            roofline["traffic_splits"] = nsplit

    eigen_outputs = None
    if mode == lam.MODE_ADAPTIVE and not args.no_eigen_outputs:
        eigen_outputs = run_eigen_outputs(shape, d_ring, d_scales, w16, args)
        log(f"eigen outputs: {eigen_outputs}")

    lambda_gpu = projection = None
    if not args.no_lambda_gpu:
        lambda_gpu = run_lambda_gpu(shape, d_ring, d_scales, w16, args)
        log(f"lambda-gpu: {lambda_gpu}")
        projection = run_projection(shape, d_ring, d_scales, args)
        log(f"projection: {projection}")

    learned = None
    if not args.no_learned:
        learned = run_learned(args)

    precision = None
    if not args.no_precision:
        precision = run_precision(args)

    subarrays = None
    if not args.no_subarrays:
        subarrays = run_subarrays(args)
        log(f"subarrays: {subarrays}")

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(shape, ring, scales, w, min(args.cpu_channels, shape.C))

    line = {
        "metric": "ant-samples/s & real-time factor, corr+beam+RFI-null, learned vs plain eig",
        "value": value,
        "unit": "ant-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8->int32 corr, fp16 MFMA beams, fp32 eig",
        "data": "synthetic (seeded Gaussian sigma=16 + 1 CW tone/channel, int8)",
        "config": {"workload": f"{args.mode}: {shape.C} ch x 2 pol x 36 ant (64 rx), "
                               f"{shape.T} samples/block, {shape.B} beams -> {nb_out} out, "
                               "plain eigen solve without eigen outputs (Householder tridiagonal, "
                               "top eigenvalues by multisection, removed vector by twisted factorization) "
                               "rank-1 null, "
                               f"int64 vis accumulated every block (dumped outside the timed region), {args.buffers} pipeline buffers "
                               "(consecutive blocks overlap)",
                   "channels": shape.C, "samples_per_block": shape.T, "beams": shape.B,
                   "eig": "plain", "buffers": args.buffers},
        "real_time_factor": rtf,
        "real_time_factor_ref_convention": (shape.T / REF_RATE_HZ) / sec,
        "null_basis_orthogonality_fp64_max": orth_max,
        "roofline": roofline,
        "eig_roofline": eig_roof,
        "with_eigen_outputs": eigen_outputs,
        "kernels": kstats,
        "autotune": tuned,
        "learned_vs_plain": learned,
        "precision_sweep": precision,
        "host_fed_pcie_inclusive": host_fed,
        "spectra_on": spectra,
        "lambda_gpu_mode": lambda_gpu,
        "projection_mode": projection,
        "uneven_subarrays": subarrays,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
