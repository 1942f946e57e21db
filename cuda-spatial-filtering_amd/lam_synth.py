"""Seeded synthetic LAMBDA voltage blocks (numpy) for tests and benchmarks.

Produces blocks in the reference's ring layout
``int8 [C][NP+2][F][NT][RPP][P][2]`` with ``int16 [C][NP+2][R][P]`` scales
(packet_formats.hpp:305-312), the shape a reassembled ``FinalPacketData``
hands to ``GPUPipeline::execute_pipeline`` (pipeline_base.hpp:28-29).

Model: i.i.d. complex Gaussian noise (sigma per component) plus CW tones
arriving as plane waves over a random 2-D antenna layout inside a 30 m
radius: tone_k(t, a) = A_k exp(i(2 pi nu_k t + phi_k)) exp(-2 pi i f_c tau_ka),
tau_ka = (l_k e_a + m_k n_a) / c.  re, im = clip(round(noise + tones), +-127).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0
CHANNEL_BW_HZ = 781250.0


@dataclass(frozen=True)
class BlockShape:
    """Runtime equivalent of LambdaConfig (packet_formats.hpp:265-407)."""

    C: int
    F: int
    NT: int
    R: int
    P: int
    RPP: int
    NP: int
    B: int
    R_pad: int = 64
    blocks_to_accumulate: int = 1

    @property
    def T(self) -> int:
        return self.NP * self.NT

    @property
    def nbl(self) -> int:
        return self.R * (self.R + 1) // 2

    def ring_shape(self):
        return (self.C, self.NP + 2, self.F, self.NT, self.RPP, self.P, 2)

    def scale_shape(self):
        return (self.C, self.NP + 2, self.R, self.P)


# Shapes of the benchmark configurations (36 antennas padded to 64 receivers).
CFG1 = BlockShape(C=8, F=1, NT=64, R=36, P=2, RPP=36, NP=16, B=4)
CFG2 = BlockShape(C=384, F=1, NT=64, R=36, P=2, RPP=36, NP=128, B=16, blocks_to_accumulate=4)


@dataclass
class Scene:
    """Geometry and interferers of one synthetic observation."""

    east: np.ndarray            # [R] metres
    north: np.ndarray           # [R]
    freqs_hz: np.ndarray        # [C] channel centre frequencies
    amps: np.ndarray            # [C][K] tone amplitudes (0 = off)
    nus: np.ndarray             # [C][K] cycles / sample
    phases: np.ndarray          # [C][K][P]
    lm: np.ndarray              # [C][K][2] direction cosines
    steering: np.ndarray = field(default=None)  # [C][K][R] complex

    def compute_steering(self):
        tau = (self.lm[..., 0, None] * self.east[None, None, :]
               + self.lm[..., 1, None] * self.north[None, None, :]) / C_LIGHT
        self.steering = np.exp(-2j * np.pi * self.freqs_hz[:, None, None] * tau)
        return self


def random_layout(rng, R, radius=30.0):
    r = radius * np.sqrt(rng.random(R))
    th = 2 * np.pi * rng.random(R)
    return r * np.cos(th), r * np.sin(th)


def random_lm(rng, shape):
    r = np.sqrt(rng.random(shape))
    th = 2 * np.pi * rng.random(shape)
    return np.stack([r * np.cos(th), r * np.sin(th)], -1)


def make_scene(shape: BlockShape, seed: int, n_tones=1, amp=40.0, inr_db=None,
               first_channel=64, n_active=None) -> Scene:
    """Scene with ``n_tones`` CW interferers per channel.

    ``amp`` fixes the amplitude; ``inr_db=(lo, hi)`` instead draws an INR per
    tone and sets A = 16*sqrt(2)*10^(INR/20)."""
    rng = np.random.default_rng(seed)
    R = shape.R if n_active is None else n_active
    e, n = random_layout(rng, R)
    if R < shape.R:
        e = np.concatenate([e, np.zeros(shape.R - R)])
        n = np.concatenate([n, np.zeros(shape.R - R)])
    freqs = (first_channel + np.arange(shape.C)) * CHANNEL_BW_HZ
    K = n_tones
    if inr_db is None:
        amps = np.full((shape.C, K), float(amp))
    else:
        inr = rng.uniform(inr_db[0], inr_db[1], (shape.C, K))
        amps = 16.0 * np.sqrt(2.0) * 10 ** (inr / 20.0)
    nus = rng.uniform(-0.5, 0.5, (shape.C, K))
    phases = rng.uniform(0, 2 * np.pi, (shape.C, K, shape.P))
    lm = random_lm(rng, (shape.C, K))
    return Scene(e, n, freqs, amps, nus, phases, lm).compute_steering()


def make_block(shape: BlockShape, scene: Scene, seed: int, sigma=16.0,
               time_offset=0, active=None):
    """Ring-layout int8 samples and unit int16 scales for one block.

    ``active``: optional [R] bool mask; inactive receivers carry zeros."""
    rng = np.random.default_rng(seed)
    C, F, NT, R, P, RPP, NP = shape.C, shape.F, shape.NT, shape.R, shape.P, shape.RPP, shape.NP
    G = (NP + 2) * NT
    ring = np.empty(shape.ring_shape(), np.int8)
    g = np.arange(G, dtype=np.float64) + time_offset
    for c in range(C):
        noise = rng.standard_normal((G, R, P, 2), dtype=np.float32) * np.float32(sigma)
        sig = np.zeros((G, R, P), np.complex128)
        for k in range(scene.amps.shape[1]):
            a = scene.amps[c, k]
            if a == 0.0:
                continue
            tone = a * np.exp(1j * (2 * np.pi * scene.nus[c, k] * g))  # [G]
            pol = np.exp(1j * scene.phases[c, k])  # [P]
            sig += tone[:, None, None] * scene.steering[c, k][None, :, None] * pol[None, None, :]
        v = noise
        v[..., 0] += sig.real.astype(np.float32)
        v[..., 1] += sig.imag.astype(np.float32)
        q = np.clip(np.rint(v), -127, 127).astype(np.int8)  # [G][R][P][2]
        if active is not None:
            q[:, ~np.asarray(active, bool)] = 0
        # [G][R][P][2] -> [NP+2][NT][F][RPP][P][2] -> [NP+2][F][NT][RPP][P][2]
        q = q.reshape(NP + 2, NT, F, RPP, P, 2).transpose(0, 2, 1, 3, 4, 5)
        ring[c] = q
    scales = np.ones(shape.scale_shape(), np.int16)
    return ring, scales


def random_weights(shape: BlockShape, seed: int, nb=None):
    """FP16 unit-modulus weights of random phase, complex<half>[C][P][B][R]
    as uint16 pairs plus their exact complex128 values."""
    rng = np.random.default_rng(seed)
    nb = shape.B if nb is None else nb
    th = rng.uniform(0, 2 * np.pi, (shape.C, shape.P, nb, shape.R))
    re = np.cos(th).astype(np.float16)
    im = np.sin(th).astype(np.float16)
    halfs = np.stack([re, im], -1)
    exact = re.astype(np.float64) + 1j * im.astype(np.float64)
    return halfs.view(np.uint16), exact


def unity_weights(shape: BlockShape, value=1.0):
    """make_unity_beam_weights (tests/support/pipeline_harness.hpp:34-44)."""
    halfs = np.zeros((shape.C, shape.P, shape.B, shape.R, 2), np.float16)
    halfs[..., 0] = value
    return halfs.view(np.uint16), halfs[..., 0].astype(np.float64) + 0j


def ring_from_fn(shape: BlockShape, sample_fn, scale_fn=None):
    """Ring block from per-sample callables, mirroring the reference harness
    (tests/support/pipeline_harness.hpp:204-221): packet k of the window is
    written to slot k+1, packet -1 to slot 0.

    sample_fn(channel, fpga, packet, time, receiver_in_packet, pol) -> (re, im)
    scale_fn(channel, fpga, packet, receiver_in_packet, pol) -> int"""
    ring = np.zeros(shape.ring_shape(), np.int8)
    scales = np.zeros(shape.scale_shape(), np.int16)
    for c in range(shape.C):
        for f in range(shape.F):
            for slot in range(shape.NP + 2):
                pkt = slot - 1
                for n in range(shape.RPP):
                    for p in range(shape.P):
                        scales[c, slot, f * shape.RPP + n, p] = (
                            1 if scale_fn is None else scale_fn(c, f, pkt, n, p))
                        for t in range(shape.NT):
                            re, im = sample_fn(c, f, pkt, t, n, p)
                            ring[c, slot, f, t, n, p, 0] = re
                            ring[c, slot, f, t, n, p, 1] = im
    return ring, scales


# ---- device generator tables (lam_synth_block) -----------------------------

def tone_tables(steer, amps, nus, phases, pol0, pol1, active):
    """Pack per-channel tone slots for lam_synth_block.

    steer [C][S][R] complex, amps/nus/phases/pol0 [C][S] real, pol1 [C][S]
    complex, active [C][S] bool.  Returns (tones float32 [C][S][6],
    steer float32 [C][S][R][2])."""
    C, S = amps.shape
    tones = np.zeros((C, S, 6), np.float32)
    tones[..., 0] = np.where(active, amps, 0.0)
    tones[..., 1] = nus
    tones[..., 2] = phases
    tones[..., 3] = pol0
    tones[..., 4] = np.real(pol1)
    tones[..., 5] = np.imag(pol1)
    st = np.stack([steer.real, steer.imag], -1).astype(np.float32)
    return tones, np.ascontiguousarray(st)


class DriftScenario:
    """Interferers whose directions drift and which switch on and off.

    Per channel 1..4 CW sources (n_src), INR drawn in inr_db with
    A = 16 sqrt(2) 10^(INR/20) (noise sigma 16 per component), azimuth
    drifting by `drift` rad per block at fixed distance from zenith on the
    (l, m) disk, and each source toggling on/off with probability p_switch
    per block.  A per-channel requantisation gain keeps noise plus tones
    inside +-127 (the INR is unchanged by it)."""

    def __init__(self, C, R, seed, n_blocks, n_src=(1, 4), inr_db=(5.0, 30.0), drift=0.01,
                 p_switch=0.1, first_channel=64, sigma=16.0, max_sources=4):
        rng = np.random.default_rng(seed)
        self.C, self.R, self.S, self.sigma = C, R, max_sources, sigma
        self.east, self.north = random_layout(rng, R)
        self.freqs = (first_channel + np.arange(C)) * CHANNEL_BW_HZ
        self.nsrc = rng.integers(n_src[0], n_src[1] + 1, C)
        S = max_sources
        self.exists = np.arange(S)[None, :] < self.nsrc[:, None]
        self.inr = rng.uniform(inr_db[0], inr_db[1], (C, S))
        self.amps = 16.0 * np.sqrt(2.0) * 10 ** (self.inr / 20.0)
        self.nus = rng.uniform(-0.5, 0.5, (C, S))
        self.phases = rng.uniform(0, 2 * np.pi, (C, S))
        psi = rng.uniform(np.pi / 8, 3 * np.pi / 8, (C, S))
        chi = rng.uniform(0, 2 * np.pi, (C, S))
        self.pol0 = np.cos(psi)
        self.pol1 = np.sin(psi) * np.exp(1j * chi)
        self.rho = np.sqrt(rng.uniform(0.0, 0.95, (C, S)))
        self.az0 = rng.uniform(0, 2 * np.pi, (C, S))
        self.drift = drift
        on = np.ones((C, S), bool)
        self.on = np.zeros((n_blocks, C, S), bool)
        for b in range(n_blocks):
            if b > 0:
                on = on ^ (rng.random((C, S)) < p_switch)
            self.on[b] = on & self.exists

    def steering(self, b):
        az = self.az0 + self.drift * b
        l, m = self.rho * np.cos(az), self.rho * np.sin(az)
        tau = (l[..., None] * self.east + m[..., None] * self.north) / C_LIGHT
        return np.exp(-2j * np.pi * self.freqs[:, None, None] * tau)  # [C][S][R]

    def block(self, b):
        """(tones, steer, gain, steer_complex, active) of block b."""
        st = self.steering(b)
        active = self.on[b]
        tones, st32 = tone_tables(st, self.amps, self.nus, self.phases, self.pol0, self.pol1, active)
        peak = 3.0 * self.sigma + np.sum(np.where(active, self.amps, 0.0), axis=1)
        gain = np.minimum(1.0, 120.0 / peak).astype(np.float32)
        return tones, st32, gain, st, active


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
    return z ^ (z >> np.uint64(31))


def _gauss2(h):
    u1 = ((h >> np.uint64(40)).astype(np.float32) + np.float32(1)) * np.float32(1 / 16777216)
    u2 = ((h >> np.uint64(8)) & np.uint64(0xffffff)).astype(np.float32) * np.float32(1 / 16777216)
    r = np.sqrt(np.float32(-2) * np.log(u1))
    a = np.float32(2 * np.pi) * u2
    return r * np.cos(a), r * np.sin(a)


def host_synth(shape: BlockShape, seed, block_index, sigma, tones, steer, gain):
    """numpy restatement of lam_synth_block (test reference; small shapes)."""
    C, F, NT, RPP, NP = shape.C, shape.F, shape.NT, shape.RPP, shape.NP
    NS = NP + 2
    S = tones.shape[1]
    with np.errstate(over="ignore"):
        gid = np.arange(C * NS * F * NT, dtype=np.uint64).reshape(C, NS, F, NT)
        base = (np.uint64(seed) * np.uint64(0x9e3779b97f4a7c15)) ^ (np.uint64(block_index) << np.uint64(44))
        rr = np.arange(RPP, dtype=np.uint64)
        ctr = (gid[..., None] * np.uint64(RPP) + rr) * np.uint64(2)  # [C][NS][F][NT][RPP]
        n0 = _gauss2(_mix64(base + ctr))
        n1 = _gauss2(_mix64(base + ctr + np.uint64(1)))
    slot = np.arange(NS)[:, None, None]
    t = np.arange(NT)[None, None, :]
    tt = ((slot - 1) * NT + t).astype(np.float64)  # [NS][1][NT]
    x0 = (sigma * n0[0] + 1j * sigma * n0[1]).astype(np.complex128)
    x1 = (sigma * n1[0] + 1j * sigma * n1[1]).astype(np.complex128)
    st = steer[..., 0] + 1j * steer[..., 1]  # [C][S][R]
    for c in range(C):
        for s in range(S):
            amp, nu, ph, g0, g1r, g1i = (float(v) for v in tones[c, s])
            if amp == 0.0:
                continue
            cyc = nu * tt
            fr = (cyc - np.floor(cyc)).astype(np.float32)
            ang = (np.float32(6.283185307179586) * fr + np.float32(ph)).astype(np.float32)
            tone = amp * (np.cos(ang) + 1j * np.sin(ang))  # [NS][1][NT]
            a = st[c, s].reshape(F, RPP)  # receiver r = f*RPP + rr
            v = tone[:, :, :, None] * a[None, :, None, :]  # [NS][F][NT][RPP]
            x0[c] += g0 * v
            x1[c] += (g1r + 1j * g1i) * v
    out = np.empty(shape.ring_shape(), np.int8)
    gb = gain.reshape(C, 1, 1, 1, 1).astype(np.float32)
    for p, x in enumerate((x0, x1)):
        xr = x.real.astype(np.float32) * gb
        xi = x.imag.astype(np.float32) * gb
        out[..., p, 0] = np.clip(np.rint(xr), -127, 127)
        out[..., p, 1] = np.clip(np.rint(xi), -127, 127)
    return out, np.ones(shape.scale_shape(), np.int16)


def null_depth_db(W, U, K, steer_active):
    """10 log10(|W P a|^2 / |W a|^2) summed over beams and active sources,
    P = I - U_K U_K^H with U the unwhitened unit-norm eigenvectors in
    ascending order (columns), W [B][R] complex, steer_active [S'][R]."""
    if K == 0 or len(steer_active) == 0:
        return 0.0
    Uk = U[:, -K:]
    P = np.eye(U.shape[0]) - Uk @ Uk.conj().T
    a = np.asarray(steer_active).T  # [R][S']
    num = np.sum(np.abs(W @ P @ a) ** 2)
    den = np.sum(np.abs(W @ a) ** 2)
    return 10 * np.log10(max(num, 1e-300) / den)
