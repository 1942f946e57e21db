"""ctypes/numpy front end of the CPU oracle (oracle/lam_oracle.c).

TEST INFRASTRUCTURE ONLY: the parity checker for the HIP product path and the
timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product never does.

Each wrapper restates one function of the reference hot path; the C side
carries the file:line citations (oracle/lam_oracle.h).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liblam_oracle.so")
_lib = None

WINDOW_CORR_BEAM = 0
WINDOW_ADAPTIVE = 1


def build() -> str:
    """Compile the oracle with the committed Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class OrcShape(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("C", "F", "NT", "R", "P", "RPP", "NP", "B")]


@dataclass(frozen=True)
class Shape:
    """LambdaConfig shape (packet_formats.hpp:265-407)."""

    C: int
    F: int
    NT: int
    R: int
    P: int
    RPP: int
    NP: int
    B: int

    @property
    def T(self) -> int:
        return self.NP * self.NT

    @property
    def nbl(self) -> int:
        return self.R * (self.R + 1) // 2

    def c(self) -> OrcShape:
        return OrcShape(self.C, self.F, self.NT, self.R, self.P, self.RPP, self.NP, self.B)


_p = ctypes.c_void_p


def _declare(L):
    L.orc_ingest.argtypes = [ctypes.POINTER(OrcShape), _p, _p, ctypes.c_int, _p, _p, _p, _p]
    L.orc_correlate.argtypes = [ctypes.POINTER(OrcShape), _p, _p]
    L.orc_beamform.argtypes = [ctypes.POINTER(OrcShape), _p, _p, ctypes.c_int, _p]
    L.orc_covariance.argtypes = [ctypes.c_int, ctypes.c_int, _p, ctypes.c_int, _p]
    L.orc_whiten.argtypes = [ctypes.c_int, _p, _p]
    L.orc_jacobi_eigh.argtypes = [ctypes.c_int, _p, _p, _p, ctypes.c_int, ctypes.c_double]
    L.orc_jacobi_eigh.restype = ctypes.c_int
    L.orc_percentile_linear.argtypes = [_p, ctypes.c_int, ctypes.c_double]
    L.orc_percentile_linear.restype = ctypes.c_double
    L.orc_detect_k.argtypes = [_p, ctypes.c_int, ctypes.c_double]
    L.orc_detect_k.restype = ctypes.c_int
    L.orc_unwhiten_normalize.argtypes = [ctypes.c_int, _p, _p]
    L.orc_null_projection.argtypes = [ctypes.c_int, _p, ctypes.c_int, _p]
    L.orc_shrink_projection.argtypes = [ctypes.c_int, _p, _p, ctypes.c_int, _p]
    L.orc_apply_projection.argtypes = [ctypes.c_int, ctypes.c_int, _p, _p, _p]
    L.orc_fine_delay_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _p,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, _p]
    L.orc_rfi_null_cp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _p, ctypes.c_int,
                                  _p, ctypes.c_int, ctypes.c_double, ctypes.c_int, _p, _p, _p]
    L.orc_rfi_null_cp.restype = ctypes.c_int
    L.orc_adaptive_block.argtypes = [ctypes.POINTER(OrcShape), _p, _p, _p, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_int, _p, _p, _p, _p, _p]
    L.orc_adaptive_block.restype = ctypes.c_int
    L.orc_num_threads.restype = ctypes.c_int
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_beam_spectra.argtypes = [ctypes.c_int] * 6 + [_p, _p]
    L.orc_bandpass.argtypes = [ctypes.c_int] * 3 + [_p, _p]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ----------------------------------------------------------------- pipeline
def ingest(shape: Shape, samples, scales, window=WINDOW_CORR_BEAM, delays=None,
           recv_perm=None, pol_perm=None) -> np.ndarray:
    """int8 ring block x int16 scales -> y int32 [C][T][R][P][2]."""
    samples = _c(samples, np.int8)
    scales = _c(scales, np.int16)
    y = np.empty((shape.C, shape.T, shape.R, shape.P, 2), np.int32)
    d = None if delays is None else _c(delays, np.int32)
    rp = None if recv_perm is None else _c(recv_perm, np.int32)
    pp = None if pol_perm is None else _c(pol_perm, np.int32)
    sh = shape.c()
    lib().orc_ingest(ctypes.byref(sh), _ptr(samples), _ptr(scales), window, _ptr(d), _ptr(rp),
                     _ptr(pp), _ptr(y))
    return y


def correlate(shape: Shape, y) -> np.ndarray:
    """Exact visibilities int64 [C][R(R+1)/2][P][P][2]."""
    y = _c(y, np.int32)
    vis = np.empty((shape.C, shape.nbl, shape.P, shape.P, 2), np.int64)
    sh = shape.c()
    lib().orc_correlate(ctypes.byref(sh), _ptr(y), _ptr(vis))
    return vis


def beamform(shape: Shape, y, w) -> np.ndarray:
    """FP64 beams complex128 [C][P][nb][T] from weights complex [C][P][nb][R]."""
    y = _c(y, np.int32)
    w = np.asarray(w)
    nb = w.shape[2]
    wr = _c(np.stack([w.real, w.imag], -1), np.float64)
    out = np.empty((shape.C, shape.P, nb, shape.T, 2), np.float64)
    sh = shape.c()
    lib().orc_beamform(ctypes.byref(sh), _ptr(y), _ptr(wr), nb, _ptr(out))
    return out[..., 0] + 1j * out[..., 1]


def _cplx_in(a, n):
    a = np.asarray(a, np.complex128).reshape(n, n)
    # column-major storage: element (i, j) at j*n + i  ->  store a.T row-major
    return _c(np.stack([a.T.real, a.T.imag], -1), np.float64)


def _cplx_out(buf, n):
    m = buf[..., 0] + 1j * buf[..., 1]
    return m.reshape(n, n).T.copy()


def covariance(R: int, P: int, vis_c, p: int) -> np.ndarray:
    vis_c = _c(vis_c, np.int64)
    A = np.empty((R, R, 2), np.float64)
    lib().orc_covariance(R, P, _ptr(vis_c), p, _ptr(A))
    return _cplx_out(A, R)


def whiten(A):
    n = A.shape[0]
    buf = _cplx_in(A, n)
    d = np.empty(n, np.float64)
    lib().orc_whiten(n, _ptr(buf), _ptr(d))
    return _cplx_out(buf, n), d


def jacobi_eigh(A, max_sweeps=100, tol=1e-15):
    """(eigenvalues ascending, eigenvector matrix (columns), sweeps)."""
    n = A.shape[0]
    buf = _cplx_in(A, n)
    ev = np.empty(n, np.float64)
    V = np.empty((n, n, 2), np.float64)
    sweeps = lib().orc_jacobi_eigh(n, _ptr(buf), _ptr(ev), _ptr(V), max_sweeps, tol)
    return ev, _cplx_out(V, n), sweeps


def percentile_linear(sorted_values, q: float) -> float:
    v = _c(sorted_values, np.float64)
    return lib().orc_percentile_linear(_ptr(v), v.size, q)


def detect_k(evals, delta: float) -> int:
    v = _c(evals, np.float64)
    return lib().orc_detect_k(_ptr(v), v.size, delta)


def unwhiten_normalize(V, d):
    n = V.shape[0]
    buf = _cplx_in(V, n)
    dd = _c(d, np.float64)
    lib().orc_unwhiten_normalize(n, _ptr(buf), _ptr(dd))
    return _cplx_out(buf, n)


def null_projection(V, k: int):
    n = V.shape[0]
    buf = _cplx_in(V, n)
    P = np.empty((n, n, 2), np.float64)
    lib().orc_null_projection(n, _ptr(buf), k, _ptr(P))
    return _cplx_out(P, n)


def shrink_projection(V, evals, k: int):
    n = V.shape[0]
    buf = _cplx_in(V, n)
    ev = _c(evals, np.float64)
    P = np.empty((n, n, 2), np.float64)
    lib().orc_shrink_projection(n, _ptr(buf), _ptr(ev), k, _ptr(P))
    return _cplx_out(P, n)


def apply_projection(W, P):
    W = np.asarray(W, np.complex128)
    nb, n = W.shape
    wb = _c(np.stack([W.real, W.imag], -1), np.float64)
    pb = _cplx_in(P, n)
    out = np.empty((nb, n, 2), np.float64)
    lib().orc_apply_projection(nb, n, _ptr(wb), _ptr(pb), _ptr(out))
    return out[..., 0] + 1j * out[..., 1]


def fine_delay_table(C, R, N, delays_ns, base_freq_hz, channel_bw_hz, min_freq_ch):
    d = _c(delays_ns, np.float32)
    out = np.empty((C, R, N, 2), np.float64)
    lib().orc_fine_delay_table(C, R, N, _ptr(d), base_freq_hz, channel_bw_hz, min_freq_ch,
                               _ptr(out))
    return out[..., 0] + 1j * out[..., 1]


def rfi_null_cp(R, P, vis_c, p, w_cp, fixed_k=-1, delta=3.0, shrink=False):
    """One (channel, pol) RFI-null problem -> (W2 [2B][R], evals, evecs, K)."""
    w_cp = np.asarray(w_cp, np.complex128)
    B = w_cp.shape[0]
    vis_c = _c(vis_c, np.int64)
    wb = _c(np.stack([w_cp.real, w_cp.imag], -1), np.float64)
    out_w = np.empty((2 * B, R, 2), np.float64)
    ev = np.empty(R, np.float64)
    V = np.empty((R, R, 2), np.float64)
    k = lib().orc_rfi_null_cp(R, P, B, _ptr(vis_c), p, _ptr(wb), int(fixed_k), float(delta),
                              int(bool(shrink)), _ptr(out_w), _ptr(ev), _ptr(V))
    return out_w[..., 0] + 1j * out_w[..., 1], ev, _cplx_out(V, R), k


def adaptive_block(shape: Shape, y, w, fixed_k=None, detect=False, delta=3.0, shrink=False):
    """Full adaptive block on y: returns dict(vis, beams2b, evals, evecs, counts)."""
    C, P, R, B, T = shape.C, shape.P, shape.R, shape.B, shape.T
    y = _c(y, np.int32)
    w = np.asarray(w, np.complex128)
    wb = _c(np.stack([w.real, w.imag], -1), np.float64)
    fk = _c(np.zeros((C, P)) if fixed_k is None else fixed_k, np.int32)
    vis = np.empty((C, shape.nbl, P, P, 2), np.int64)
    beams = np.empty((C, P, 2 * B, T, 2), np.float64)
    ev = np.empty((C, P, R), np.float64)
    V = np.empty((C, P, R, R, 2), np.float64)
    counts = np.empty((C, P), np.int32)
    sh = shape.c()
    lib().orc_adaptive_block(ctypes.byref(sh), _ptr(y), _ptr(wb), _ptr(fk), int(bool(detect)),
                             float(delta), int(bool(shrink)), _ptr(vis), _ptr(beams), _ptr(ev),
                             _ptr(V), _ptr(counts))
    return dict(vis=vis, beams2b=beams[..., 0] + 1j * beams[..., 1], evals=ev,
                evecs=V[..., 0] + 1j * V[..., 1], counts=counts)


def beam_spectra(beams, nfft=64, drop=5):
    """beams complex [C][P][E][T] -> float64 [C][P][E][nfft - 2*drop]: per packet
    DFT magnitudes with the edge bins dropped, DC centred, summed over packets."""
    beams = np.asarray(beams)
    C, P, E, T = beams.shape
    assert T % nfft == 0
    b = _c(np.stack([beams.real, beams.imag], -1), np.float64)
    out = np.empty((C, P, E, nfft - 2 * drop), np.float64)
    lib().orc_beam_spectra(C, P, E, T // nfft, nfft, drop, _ptr(b), _ptr(out))
    return out


def bandpass(beams, d):
    """beams complex [..., n] -> float64 [..., n / d]: mean |DFT|^2 over runs of d bins."""
    beams = np.asarray(beams)
    lead, n = beams.shape[:-1], beams.shape[-1]
    assert n % d == 0
    rows = int(np.prod(lead)) if lead else 1
    b = _c(np.stack([beams.real, beams.imag], -1).reshape(rows, n, 2), np.float64)
    out = np.empty((rows, n // d), np.float64)
    lib().orc_bandpass(rows, n, d, _ptr(b), _ptr(out))
    return out.reshape(*lead, n // d)


def num_threads() -> int:
    return lib().orc_num_threads()


def set_threads(n: int) -> None:
    lib().orc_set_threads(n)
