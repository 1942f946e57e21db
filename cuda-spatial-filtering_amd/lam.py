"""Python binding of the C ABI in include/lam.h (liblam.so, gfx950 HIP).

Host-side mirror of the reference's pipeline interface for this path
(GPUPipeline, include/spatial/pipeline_base.hpp:9-40): a ``Pipeline`` is one
lam_ctx; ``execute`` is execute_pipeline, ``dump_visibilities`` is
dump_visibilities.  PyTorch is used only for device memory in tests and
benchmarks; there is no CPU fallback: if liblam.so is missing this module
raises on import of the library.
"""
from __future__ import annotations

import collections
import ctypes
import threading
import os
import re

import numpy as np

from lam_synth import BlockShape

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LAM_LIBRARY", os.path.join(_HERE, "liblam.so"))
HEADER = os.path.join(os.path.dirname(_HERE), "include", "lam.h")

LAM_OK, LAM_ERR_ARG, LAM_ERR_HIP, LAM_ERR_RANGE, LAM_ERR_STATE, LAM_ERR_UNSUPPORTED, LAM_ERR_NO_DEVICE = range(7)
MODE_CORR_BEAM, MODE_ADAPTIVE, MODE_LAMBDA_GPU, MODE_PROJECTION = 0, 1, 2, 3
K_CORRBEAM, K_CORR, K_EIG, K_BEAM2B = range(4)
KERNEL_NAMES = ["corr_beam", "corr", "eig_null", "beam_2b", "features", "mlp", "eig_fallback",
                "subarray_eig", "spectra", "pol_eig", "bandpass"]
EIG_PLAIN, EIG_LEARNED, EIG_PREVIOUS = 0, 1, 2
WPREC_F16_HILO, WPREC_F16, WPREC_BF16 = 0, 1, 2

_lib = None


class LamError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lam error {code}: {msg}")
        self.code = code


class LamShape(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "nr_channels", "nr_fpga_sources", "nr_time_steps_per_packet", "nr_receivers",
        "nr_polarizations", "nr_receivers_per_packet", "nr_packets_for_correlation", "nr_beams",
        "nr_padded_receivers", "nr_blocks_to_accumulate")]


HOSTFN = ctypes.CFUNCTYPE(None, ctypes.c_void_p)


class LamBlockIO(ctypes.Structure):
    _fields_ = [
        ("samples", ctypes.c_void_p), ("scales", ctypes.c_void_p),
        ("inputs_on_device", ctypes.c_int32),
        ("release_fn", HOSTFN), ("release_user", ctypes.c_void_p),
        ("beams_dst", ctypes.c_void_p), ("beams_done_fn", HOSTFN), ("beams_done_user", ctypes.c_void_p),
        ("eigenvalues_dst", ctypes.c_void_p), ("eigenvectors_dst", ctypes.c_void_p),
        ("counts_dst", ctypes.c_void_p), ("eigen_done_fn", HOSTFN), ("eigen_done_user", ctypes.c_void_p),
        ("spectra_dst", ctypes.c_void_p), ("spectra_done_fn", HOSTFN), ("spectra_done_user", ctypes.c_void_p),
    ]


def header_symbols() -> list[str]:
    """Every lam_* function declared in include/lam.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lam_[a-z0-9_]+)\s*\(", text)))


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is not built; run `make -C cuda-spatial-filtering_amd` (or __graft_entry__.build())")
    L = ctypes.CDLL(path)
    vp, i32 = ctypes.c_void_p, ctypes.c_int32
    L.lam_create.argtypes = [ctypes.POINTER(LamShape), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.lam_destroy.argtypes = [vp]
    L.lam_last_error.argtypes = [vp]
    L.lam_last_error.restype = ctypes.c_char_p
    L.lam_set_weights.argtypes = [vp, vp]
    L.lam_set_gains.argtypes = [vp, vp]
    L.lam_set_subpacket_delays.argtypes = [vp, vp]
    L.lam_set_fine_delays.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, i32]
    L.lam_set_stream_permutation.argtypes = [vp, vp, vp, ctypes.c_size_t]
    L.lam_set_fixed_k.argtypes = [vp, vp]
    L.lam_set_nulling.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_int]
    L.lam_set_spectra.argtypes = [vp, ctypes.c_int]
    L.lam_set_bandpass.argtypes = [vp, ctypes.c_int]
    L.lam_set_whitening.argtypes = [vp, ctypes.c_int]
    L.lam_set_projection.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.lam_projection_runs.argtypes = [vp]
    L.lam_dump_projection.argtypes = [vp, vp, vp, HOSTFN, vp]
    L.lam_set_eig_params.argtypes = [vp, ctypes.c_int, ctypes.c_float]
    L.lam_set_graphs.argtypes = [vp, ctypes.c_int]
    L.lam_set_profiling.argtypes = [vp, ctypes.c_int]
    L.lam_set_tuning.argtypes = [vp, ctypes.c_int]
    L.lam_kernel_times.argtypes = [vp, vp, vp]
    L.lam_execute.argtypes = [vp, ctypes.POINTER(LamBlockIO)]
    L.lam_dump_visibilities.argtypes = [vp, vp, HOSTFN, vp]
    L.lam_dump_visibilities_i64.argtypes = [vp, vp, HOSTFN, vp]
    L.lam_blocks_accumulated.argtypes = [vp]
    L.lam_sync.argtypes = [vp]
    L.lam_next_stream.argtypes = [vp]
    L.lam_next_stream.restype = vp
    L.lam_device_alloc.argtypes = [ctypes.c_size_t]
    L.lam_device_alloc.restype = vp
    L.lam_device_free.argtypes = [vp]
    L.lam_memcpy.argtypes = [vp, vp, ctypes.c_size_t]
    L.lam_device_count.restype = ctypes.c_int
    L.lam_host_alloc.argtypes = [ctypes.c_size_t]
    L.lam_host_alloc.restype = vp
    L.lam_host_free.argtypes = [vp]
    L.lam_eigh_batched.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_float, vp, vp, vp, vp]
    L.lam_set_eig_method.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float]
    L.lam_eig_stats.argtypes = [vp, vp, vp, vp]
    L.lam_set_eig_tuning.argtypes = [vp, ctypes.c_int]
    L.lam_set_weight_precision.argtypes = [vp, ctypes.c_int]
    L.lam_set_subarrays.argtypes = [vp, ctypes.c_int, vp, vp]
    L.lam_set_persistent_workgroups.argtypes = [vp, ctypes.c_int]
    L.lam_subarray_results.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.lam_autotune.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]
    L.lam_eigh_warm.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_float, vp, vp, vp, vp]
    L.lam_learn_init.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
    L.lam_learn_dims.argtypes = [vp, vp, vp]
    L.lam_learn_capture.argtypes = [vp, ctypes.c_int]
    L.lam_learn_captured.argtypes = [vp]
    L.lam_learn_sample.argtypes = [vp, ctypes.c_int, vp, vp, vp]
    L.lam_learn_last.argtypes = [vp, vp, vp]
    L.lam_learn_train.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, vp]
    L.lam_learn_params.argtypes = [vp, vp, ctypes.c_int64, ctypes.c_int]
    L.lam_learn_grad.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp, vp]
    L.lam_graph_replays.argtypes = [vp]
    L.lam_graph_replays.restype = ctypes.c_longlong
    L.lam_correlate.argtypes = [ctypes.POINTER(LamShape), ctypes.c_int, vp, vp, vp, vp, vp]
    L.lam_correlate_noalloc.argtypes = [ctypes.POINTER(LamShape), ctypes.c_int, vp, vp, vp, vp, vp]
    L.lam_synth_block.argtypes = [ctypes.POINTER(LamShape), ctypes.c_uint64, i32, ctypes.c_float, i32,
                                  vp, vp, vp, vp, vp, vp]
    _lib = L
    return L


class DeviceArray:
    """Device buffer owned by liblam (hipMalloc through the C ABI) with numpy
    shape/dtype metadata; avoids a second HIP runtime in the process."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.L = load_library()
        self.ptr = self.L.lam_device_alloc(self.nbytes)
        if not self.ptr:
            raise LamError(LAM_ERR_HIP, f"lam_device_alloc({self.nbytes}) failed")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        if self.L.lam_memcpy(self.ptr, a.ctypes.data, self.nbytes) != LAM_OK:
            raise LamError(LAM_ERR_HIP, "upload failed")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.L.lam_memcpy(out.ctypes.data, self.ptr, self.nbytes) != LAM_OK:
            raise LamError(LAM_ERR_HIP, "download failed")
        return out

    def free(self):
        if self.ptr:
            self.L.lam_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def memcpy(dst, src, nbytes):
    """Synchronous copy between any host/device buffers (DeviceArray or address)."""
    L = load_library()
    if L.lam_memcpy(_ptr(dst), _ptr(src), nbytes) != LAM_OK:
        raise LamError(LAM_ERR_HIP, "lam_memcpy failed")


def device_count() -> int:
    return load_library().lam_device_count()


class HostArray:
    """Pinned host buffer (lam_host_alloc) viewed as a numpy array: H2D copies
    from it are DMA at full link speed and run asynchronously on the
    pipeline's streams (LambdaFinalPacketData's cudaHostAlloc buffers,
    packet_formats.hpp:131-142)."""

    def __init__(self, shape, dtype):
        self.L = load_library()
        dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(shape)) * dtype.itemsize
        self.ptr = self.L.lam_host_alloc(self.nbytes)
        if not self.ptr:
            raise LamError(LAM_ERR_HIP, f"lam_host_alloc({self.nbytes}) failed")
        raw = (ctypes.c_uint8 * self.nbytes).from_address(self.ptr)
        self.array = np.frombuffer(raw, dtype=dtype).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            self.L.lam_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostRing:
    """Pinned input blocks recycled through the pipeline's release callback:
    the buffer life cycle of the reference's ProcessorState
    (include/spatial/spatial.hpp:718-770).  acquire() waits for a buffer the
    pipeline has finished copying to the device; submit() enqueues it
    (lam_execute with release_fn), so the producer fills one buffer while
    the H2D copies and kernels of the others run."""

    def __init__(self, pipeline, num_buffers):
        self.p = pipeline
        s = pipeline.shape
        self.samples = [HostArray(s.ring_shape(), np.int8) for _ in range(num_buffers)]
        self.scales = [HostArray(s.scale_shape(), np.int16) for _ in range(num_buffers)]
        self.free = collections.deque(range(num_buffers))
        self.cv = threading.Condition()
        self.released = 0
        self._cb = HOSTFN(self._release)

    def _release(self, user):
        with self.cv:
            self.free.append(int(user or 0))
            self.released += 1
            self.cv.notify()

    def acquire(self, timeout=120.0):
        with self.cv:
            if not self.cv.wait_for(lambda: len(self.free) > 0, timeout):
                raise TimeoutError("no pinned buffer released")
            return self.free.popleft()

    def submit(self, i, **outputs):
        self.p.execute(self.samples[i].array, self.scales[i].array, on_device=False,
                       release_fn=self._cb, release_user=i, **outputs)

    def close(self):
        for a in self.samples + self.scales:
            a.free()


def _ptr(x):
    """Raw address of a numpy array (host), DeviceArray or torch tensor (device)."""
    if x is None:
        return None
    if isinstance(x, HostArray):
        return x.ptr
    if isinstance(x, np.ndarray):
        assert x.flags.c_contiguous
        return x.ctypes.data
    if isinstance(x, DeviceArray):
        return x.ptr
    assert x.is_contiguous()
    return x.data_ptr()


_NULLFN = HOSTFN()


class Pipeline:
    """One lam_ctx: LambdaCorrBeamOnlyGPUPipeline (mode 0) or
    LambdaAdaptiveBeamformedSpectraPipeline (mode 1) on the current GPU."""

    def __init__(self, shape: BlockShape, mode=MODE_CORR_BEAM, num_buffers=1):
        self.L = load_library()
        self.shape = shape
        self.mode = mode
        sh = LamShape(shape.C, shape.F, shape.NT, shape.R, shape.P, shape.RPP, shape.NP, shape.B,
                      shape.R_pad, shape.blocks_to_accumulate)
        h = ctypes.c_void_p()
        rc = self.L.lam_create(ctypes.byref(sh), mode, num_buffers, ctypes.byref(h))
        if rc != LAM_OK:
            raise LamError(rc, "lam_create failed")
        self.h = h
        self._keep = []
        self.spectra = False
        self.bandpass = 0

    # -- plumbing
    def _chk(self, rc):
        if rc != LAM_OK:
            raise LamError(rc, self.L.lam_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.lam_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_out_beams(self):
        if self.mode != MODE_ADAPTIVE:
            return self.shape.B
        return self.shape.B * (2 + getattr(self, "_sub_G", 0))

    # -- setters (GPUPipeline virtuals)
    def set_weights(self, halfs_u16):
        a = np.ascontiguousarray(halfs_u16, np.uint16)
        self._chk(self.L.lam_set_weights(self.h, a.ctypes.data))

    def set_gains(self, gains):
        g = np.ascontiguousarray(np.stack([np.real(gains), np.imag(gains)], -1), np.float32)
        self._chk(self.L.lam_set_gains(self.h, g.ctypes.data))

    def set_subpacket_delays(self, delays):
        d = np.ascontiguousarray(delays, np.int32)
        self._chk(self.L.lam_set_subpacket_delays(self.h, d.ctypes.data))

    def set_fine_delays(self, delays_ns, base_freq_hz, channel_bw_hz, min_freq_ch):
        d = None if delays_ns is None else np.ascontiguousarray(delays_ns, np.float32)
        self._chk(self.L.lam_set_fine_delays(self.h, None if d is None else d.ctypes.data,
                                             base_freq_hz, channel_bw_hz, min_freq_ch))

    def set_stream_permutation(self, recv_perm, pol_perm):
        r = np.ascontiguousarray(recv_perm, np.int32)
        p = np.ascontiguousarray(pol_perm, np.int32)
        self._chk(self.L.lam_set_stream_permutation(self.h, r.ctypes.data, p.ctypes.data, r.size))

    def set_fixed_k(self, k):
        a = np.ascontiguousarray(np.broadcast_to(k, (self.shape.C, 2)), np.int32)
        self._chk(self.L.lam_set_fixed_k(self.h, a.ctypes.data))

    def set_nulling(self, detect=False, delta=3.0, shrink=False):
        self._chk(self.L.lam_set_nulling(self.h, int(bool(detect)), float(delta), int(bool(shrink))))

    def set_spectra(self, enable=True):
        """Post-beam spectra (adaptive mode, 64 time steps per packet)."""
        self._chk(self.L.lam_set_spectra(self.h, int(bool(enable))))
        self.spectra = bool(enable)

    def set_whitening(self, enable=True):
        """Off: the pulsar-fold RFI branch (eigenvectors of the raw covariance)."""
        self._chk(self.L.lam_set_whitening(self.h, int(bool(enable))))

    def set_projection(self, nr_signal_eigenvectors, nr_runs_to_average):
        self._chk(self.L.lam_set_projection(self.h, int(nr_signal_eigenvectors), int(nr_runs_to_average)))

    def projection_runs(self):
        return int(self.L.lam_projection_runs(self.h))

    def dump_projection(self, wait=True, out=None):
        """Average of the accumulated projectors, decomposed: (evals [C][2][2][R],
        evecs complex [C][2][2][R][R], eigenvector k at [..., k, :]).  wait=False
        only enqueues the dump (the arrays fill once the stream gets there) and
        returns the raw float32 landing arrays; out = (evals, evecs) arrays of
        those shapes (pinned HostArray views keep the copies asynchronous)."""
        s = self.shape
        if out is None:
            ev = np.empty((s.C, 2, 2, s.R), np.float32)
            vec = np.empty((s.C, 2, 2, s.R, s.R, 2), np.float32)
        else:
            ev, vec = out
        self._chk(self.L.lam_dump_projection(self.h, ev.ctypes.data, vec.ctypes.data, _NULLFN, None))
        if not wait:
            if out is None:  # our own arrays must outlive the copy; the caller keeps theirs
                self._keep.append((ev, vec))
            return ev, vec
        self.sync()
        self._keep.clear()
        return ev, vec[..., 0] + 1j * vec[..., 1]

    def set_bandpass(self, downsample_factor):
        """LambdaGPU bandpass: mean |FFT|^2 over runs of downsample_factor bins (0: off)."""
        self._chk(self.L.lam_set_bandpass(self.h, int(downsample_factor)))
        self.bandpass = int(downsample_factor)

    def spectra_array(self):
        s = self.shape
        if self.mode == MODE_LAMBDA_GPU:
            return np.zeros((s.C, 2, s.B, s.T // self.bandpass), np.float32)
        return np.zeros((s.C, 2, 2 * s.B, s.NT - 10), np.float32)

    def set_eig_params(self, max_sweeps=12, tol=1e-6):
        self._chk(self.L.lam_set_eig_params(self.h, int(max_sweeps), float(tol)))

    def graph_replays(self):
        return int(self.L.lam_graph_replays(self.h))

    def set_graphs(self, enable=True):
        self._chk(self.L.lam_set_graphs(self.h, int(bool(enable))))

    def set_profiling(self, enable=True):
        self._chk(self.L.lam_set_profiling(self.h, int(bool(enable))))

    def set_tuning(self, samples_per_wg=0):
        self._chk(self.L.lam_set_tuning(self.h, int(samples_per_wg)))

    # -- per block
    def execute(self, samples, scales, beams_out=None, evals_out=None, evecs_out=None,
                counts_out=None, on_device=None, spectra_out=None, release_fn=None, release_user=None):
        """Enqueue one block.  samples/scales: numpy (host) or torch (device)."""
        if on_device is None:
            on_device = not isinstance(samples, np.ndarray)
        io = LamBlockIO()
        io.samples = _ptr(samples)
        io.scales = _ptr(scales)
        io.inputs_on_device = int(on_device)
        io.release_fn = _NULLFN if release_fn is None else release_fn
        io.release_user = release_user
        io.beams_dst = _ptr(beams_out)
        io.beams_done_fn = _NULLFN
        io.eigenvalues_dst = _ptr(evals_out)
        io.eigenvectors_dst = _ptr(evecs_out)
        io.counts_dst = _ptr(counts_out)
        io.eigen_done_fn = _NULLFN
        io.spectra_dst = _ptr(spectra_out)
        io.spectra_done_fn = _NULLFN
        self._chk(self.L.lam_execute(self.h, ctypes.byref(io)))

    def sync(self):
        return self._chk(self.L.lam_sync(self.h))

    def dump_visibilities(self):
        s = self.shape
        out = np.empty((s.C, s.nbl, 2, 2, 2), np.float32)
        self._chk(self.L.lam_dump_visibilities(self.h, out.ctypes.data, _NULLFN, None))
        self.sync()
        return out

    def dump_visibilities_i64(self):
        s = self.shape
        out = np.empty((s.C, s.nbl, 2, 2, 2), np.int64)
        self._chk(self.L.lam_dump_visibilities_i64(self.h, out.ctypes.data, _NULLFN, None))
        self.sync()
        return out

    def blocks_accumulated(self):
        return self.L.lam_blocks_accumulated(self.h)

    def next_stream(self):
        return self.L.lam_next_stream(self.h)

    def kernel_times(self):
        k = len(KERNEL_NAMES)
        ms = np.zeros(k, np.float64)
        n = np.zeros(k, np.int32)
        self._chk(self.L.lam_kernel_times(self.h, ms.ctypes.data, n.ctypes.data))
        return {KERNEL_NAMES[i]: (float(ms[i]), int(n[i])) for i in range(k) if n[i]}

    # -- eigen-stage method and the learned warm start
    def set_eig_method(self, method, refine_sweeps=2, power=False, residual_tol=0.1):
        self._chk(self.L.lam_set_eig_method(self.h, method, refine_sweeps, int(power), residual_tol))

    def autotune(self, samples_list, scales_list, iters=3):
        """Tune on device-resident blocks; returns the chosen settings."""
        n = len(samples_list)
        sp = (ctypes.c_void_p * n)(*[_ptr(x) for x in samples_list])
        sc = (ctypes.c_void_p * n)(*[_ptr(x) for x in scales_list])
        res = np.zeros(3, np.int32)
        self._chk(self.L.lam_autotune(self.h, sp, sc, n, iters, res.ctypes.data))
        return dict(samples_per_wg=int(res[0]), plain_threads=int(res[1]), refine_steps=int(res[2]))

    def set_subarrays(self, mask, k):
        """mask [C][G][R] bool/uint8, k [C][2][G] int."""
        m = np.ascontiguousarray(mask, np.uint8)
        kk = np.ascontiguousarray(k, np.int32)
        self._chk(self.L.lam_set_subarrays(self.h, m.shape[1], m.ctypes.data, kk.ctypes.data))
        self._sub_G = m.shape[1]

    def clear_subarrays(self):
        self._chk(self.L.lam_set_subarrays(self.h, 0, None, None))
        self._sub_G = 0

    def set_persistent_workgroups(self, max_workgroups=0):
        self._chk(self.L.lam_set_persistent_workgroups(self.h, int(max_workgroups)))

    def subarray_results(self):
        s = self.shape
        npb = s.C * 2 * self._sub_G
        ev = np.zeros((npb, s.R), np.float32)
        vec = np.zeros((npb, s.R * s.R, 2), np.float32)
        pr = np.zeros((npb, s.R * s.R, 2), np.float32)
        cnt = np.zeros(npb, np.int32)
        sz = np.zeros(npb, np.int32)
        sw = np.zeros(npb, np.int32)
        self._chk(self.L.lam_subarray_results(self.h, ev.ctypes.data, vec.ctypes.data, pr.ctypes.data,
                                              cnt.ctypes.data, sz.ctypes.data, sw.ctypes.data))
        return dict(evals=ev, evecs=vec[..., 0] + 1j * vec[..., 1], proj=pr[..., 0] + 1j * pr[..., 1],
                    counts=cnt, sizes=sz, sweeps=sw)

    def set_weight_precision(self, precision):
        self._chk(self.L.lam_set_weight_precision(self.h, precision))

    def set_eig_tuning(self, plain_threads=256):
        self._chk(self.L.lam_set_eig_tuning(self.h, plain_threads))

    def eig_stats(self, with_orthogonality=False):
        s = self.shape
        sw = np.zeros((s.C, 2), np.int32)
        fb = np.zeros((s.C, 2), np.int32)
        orth = np.zeros((s.C, 2), np.float32)
        self._chk(self.L.lam_eig_stats(self.h, sw.ctypes.data, fb.ctypes.data, orth.ctypes.data))
        return (sw, fb, orth) if with_orthogonality else (sw, fb)

    def learn_init(self, warm_vectors=4, hidden=256, seed=1):
        self._chk(self.L.lam_learn_init(self.h, warm_vectors, hidden, seed))

    def learn_dims(self):
        d = np.zeros(7, np.int32)
        P = ctypes.c_int64()
        self._chk(self.L.lam_learn_dims(self.h, d.ctypes.data, ctypes.byref(P)))
        keys = ["KW", "n", "Dx", "Dp", "H", "Dy", "Dyp"]
        out = {k: int(v) for k, v in zip(keys, d)}
        out["P"] = int(P.value)
        return out

    def learn_capture(self, max_blocks):
        self._chk(self.L.lam_learn_capture(self.h, max_blocks))

    def learn_captured(self):
        return int(self.L.lam_learn_captured(self.h))

    def learn_last(self):
        """Features (BF16 bits) and network output of the last learned block."""
        d = self.learn_dims()
        M = self.shape.C * 2
        x = np.zeros((M, d["Dp"]), np.uint16)
        y = np.zeros((M, d["Dyp"]), np.float32)
        self._chk(self.L.lam_learn_last(self.h, x.ctypes.data, y.ctypes.data))
        return x, y

    def learn_sample(self, block):
        """Captured features (BF16 bits), labels and mode counts of one block."""
        d = self.learn_dims()
        M = self.shape.C * 2
        x = np.zeros((M, d["Dp"]), np.uint16)
        lab = np.zeros((M, d["KW"], self.shape.R, 2), np.float32)
        k = np.zeros(M, np.int32)
        self._chk(self.L.lam_learn_sample(self.h, int(block), x.ctypes.data, lab.ctypes.data, k.ctypes.data))
        return x, lab, k

    def learn_train(self, epochs, batch=768, lr=1e-3, seed=1):
        loss = np.zeros(max(1, epochs), np.float32)
        self._chk(self.L.lam_learn_train(self.h, epochs, batch, lr, seed, loss.ctypes.data))
        return loss[:epochs]

    def learn_params(self, new=None):
        P = self.learn_dims()["P"]
        if new is None:
            out = np.zeros(P, np.float32)
            self._chk(self.L.lam_learn_params(self.h, out.ctypes.data, P, 0))
            return out
        arr = np.ascontiguousarray(new, np.float32)
        assert arr.size == P
        self._chk(self.L.lam_learn_params(self.h, arr.ctypes.data, P, 1))

    def learn_grad(self, x_bf16, labels, k):
        x = np.ascontiguousarray(x_bf16, np.uint16)
        u = np.ascontiguousarray(labels, np.float32)
        kk = np.ascontiguousarray(k, np.int32)
        P = self.learn_dims()["P"]
        g = np.zeros(P, np.float32)
        loss = ctypes.c_float()
        self._chk(self.L.lam_learn_grad(self.h, x.ctypes.data, u.ctypes.data, kk.ctypes.data, len(kk),
                                        ctypes.byref(loss), g.ctypes.data))
        return float(loss.value), g

    # -- output helpers
    def beams_array(self):
        s = self.shape
        return np.zeros((s.C, 2, self.n_out_beams, s.T, 2), np.float16)

    def run_host(self, ring, scales, eigen=True):
        """Convenience: one block from host arrays, returns fp16 beams (+eigen).

        eigen=False (adaptive mode): only the mode counts are requested, so
        no eigenvalues or eigenvectors land anywhere."""
        if self.mode == MODE_PROJECTION:  # no per-block product: only the accumulator
            self.execute(np.ascontiguousarray(ring, np.int8), np.ascontiguousarray(scales, np.int16))
            self.sync()
            return {}
        beams = self.beams_array()
        s = self.shape
        extra = {}
        if self.mode == MODE_ADAPTIVE:
            extra = dict(evals_out=np.zeros((s.C, 2, s.R), np.float32),
                         evecs_out=np.zeros((s.C, 2, s.R, s.R, 2), np.float32),
                         counts_out=np.zeros((s.C, 2), np.int32))
            if not eigen:
                del extra["evals_out"], extra["evecs_out"]
            if self.spectra:
                extra["spectra_out"] = self.spectra_array()
        elif self.mode == MODE_LAMBDA_GPU:
            extra = dict(evals_out=np.zeros((s.C, 2, 2, s.R), np.float32),
                         evecs_out=np.zeros((s.C, 2, 2, s.R, s.R, 2), np.float32))
            if self.bandpass:
                extra["spectra_out"] = self.spectra_array()
        self.execute(np.ascontiguousarray(ring, np.int8), np.ascontiguousarray(scales, np.int16),
                     beams_out=beams, **extra)
        self.sync()
        out = dict(beams=beams[..., 0].astype(np.float32) + 1j * beams[..., 1].astype(np.float32))
        out.update(extra)
        return out


def synth_block(shape, seed, block_index, sigma, tones_dev, steer_dev, gain_dev, n_sources,
                samples_dev, scales_dev, stream=None):
    """Seeded synthetic ring on the device (lam_synth_block)."""
    L = load_library()
    sh = LamShape(shape.C, shape.F, shape.NT, shape.R, shape.P, shape.RPP, shape.NP, shape.B,
                  shape.R_pad, shape.blocks_to_accumulate)
    rc = L.lam_synth_block(ctypes.byref(sh), seed, block_index, sigma, n_sources, _ptr(tones_dev),
                           _ptr(steer_dev), _ptr(gain_dev), _ptr(samples_dev), _ptr(scales_dev), stream)
    if rc != LAM_OK:
        raise LamError(rc, "lam_synth_block failed")


def eigh_warm(A_dev, prop_dev, n, nmat, warm_vectors, k, steps, tol, evals_dev, evecs_dev,
              ratio_dev, stream=None):
    L = load_library()
    rc = L.lam_eigh_warm(_ptr(A_dev), _ptr(prop_dev), n, nmat, warm_vectors, k, steps, tol,
                         _ptr(evals_dev), _ptr(evecs_dev), _ptr(ratio_dev), stream)
    if rc != LAM_OK:
        raise LamError(rc, "lam_eigh_warm failed")


def eigh_batched(A_dev, n, nmat, evals_dev, evecs_dev, sweeps_dev=None, whiten=False,
                 max_sweeps=12, tol=1e-6, stream=None):
    L = load_library()
    rc = L.lam_eigh_batched(_ptr(A_dev), n, nmat, int(whiten), max_sweeps, tol, _ptr(evals_dev),
                            _ptr(evecs_dev), _ptr(sweeps_dev), stream)
    if rc != LAM_OK:
        raise LamError(rc, "lam_eigh_batched failed")


def correlate(shape, window, samples_dev, scales_dev, vis_dev, flags_dev=None, stream=None):
    """lam_correlate: add one device block's exact int64 visibilities into vis_dev."""
    L = load_library()
    sh = LamShape(shape.C, shape.F, shape.NT, shape.R, shape.P, shape.RPP, shape.NP, shape.B,
                  shape.R_pad, shape.blocks_to_accumulate)
    rc = L.lam_correlate(ctypes.byref(sh), int(window), _ptr(samples_dev), _ptr(scales_dev), _ptr(vis_dev),
                         _ptr(flags_dev), stream)
    if rc != LAM_OK:
        raise LamError(rc, "lam_correlate failed")


def correlate_noalloc(shape, window, samples_dev, scales_dev, vis_dev, flags_dev=None, stream=None):
    """lam_correlate_noalloc: as correlate(), with the library's static scratch
    (no device allocation; enqueues on `stream` and returns)."""
    L = load_library()
    sh = LamShape(shape.C, shape.F, shape.NT, shape.R, shape.P, shape.RPP, shape.NP, shape.B,
                  shape.R_pad, shape.blocks_to_accumulate)
    rc = L.lam_correlate_noalloc(ctypes.byref(sh), int(window), _ptr(samples_dev), _ptr(scales_dev),
                                 _ptr(vis_dev), _ptr(flags_dev), stream)
    if rc != LAM_OK:
        raise LamError(rc, "lam_correlate_noalloc failed")
