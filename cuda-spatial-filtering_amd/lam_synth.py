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
