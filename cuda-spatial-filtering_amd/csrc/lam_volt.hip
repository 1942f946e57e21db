// Voltage kernel: fused ingest -> int8-MFMA correlator -> FP16-MFMA beamformer.
//
// Replaces, for one block, the reference chain
//   scale_and_convert_to_half (spatial.cuh:420-486) -> packet_to_corr_input
//   (spatial.cuh:215-304) -> TCC correlator (lambda_corr_beam_only_gpu_pipeline.hpp:314-316)
//   -> accumulate_visibilities_from_corr (src/spatial.cu:146-170)
//   -> packet_to_col_maj_cons (spatial.cuh:306-374) -> ccglib beam GEMM
//   (lambda_corr_beam_only_gpu_pipeline.hpp:335-338) -> beam_ccglib_to_half_output
//   (spatial.cuh:376-418)
// and the adaptive pipeline's apply_delays / reorder_streams
// (spatial.cuh:1420-1506, 593-649) with a single pass over the int8 voltages.
//
// Workgroup = (channel, time split).  Each 64-sample chunk of the ring block is
// copied into LDS as [time][antenna] dwords (one dword = 2 pols x re/im int8),
// with the int16 scale, FPGA delay and stream permutation applied on the way.
//
// Correlator: real rows are (antenna, component v = 2*pol + re/im).  Antennas
// form groups of 16; an MFMA tile pairs component v of antenna group A with
// component w of group B, so one v_mfma_i32_16x16x64_i8 consumes 64 time
// samples.  Per (A <= B, p, q) three int32 accumulators hold
//   re = sum Re_a Re_b + Im_a Im_b,  x = sum Im_a Re_b,  y = sum Re_a Im_b
// and V = re + i (x - y) = sum x_a conj(x_b): exact integer arithmetic, no
// negated operands (so -128 is representable).
//
// Beamformer: per pol, [Re;Im] rows of 2B beams x K = (antenna, re/im) with
// v_mfma_f32_32x32x16_f16; weights arrive pre-split into hi + lo FP16
// operands so that the product keeps ~FP32 weight precision.
#include "lam_internal.hpp"

namespace lam {

// 4x4 byte transpose: x[k] holds bytes (c0,c1,c2,c3) of time k; y[c] holds
// bytes (t0,t1,t2,t3) of component c.  v_perm_b32(hi, lo, sel) selects bytes
// of the 64-bit value {hi, lo}.
__device__ __forceinline__ void transpose4x4(uint32_t x0, uint32_t x1, uint32_t x2,
                                             uint32_t x3, uint32_t &y0, uint32_t &y1,
                                             uint32_t &y2, uint32_t &y3) {
  const uint32_t a_lo = __builtin_amdgcn_perm(x1, x0, 0x05010400u);
  const uint32_t a_hi = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
  const uint32_t b_lo = __builtin_amdgcn_perm(x3, x2, 0x05010400u);
  const uint32_t b_hi = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
  y0 = __builtin_amdgcn_perm(b_lo, a_lo, 0x05040100u);
  y1 = __builtin_amdgcn_perm(b_lo, a_lo, 0x07060302u);
  y2 = __builtin_amdgcn_perm(b_hi, a_hi, 0x05040100u);
  y3 = __builtin_amdgcn_perm(b_hi, a_hi, 0x07060302u);
}

// Multiply the two int8 (re, im) of one pol by an int16 scale; flags values
// that leave [-128, 127].
__device__ __forceinline__ uint32_t scale_pair(uint32_t pair, int s, bool &bad) {
  const int re = (int)(int8_t)(pair & 0xffu) * s;
  const int im = (int)(int8_t)((pair >> 8) & 0xffu) * s;
  bad |= (re < -128) | (re > 127) | (im < -128) | (im > 127);
  return ((uint32_t)re & 0xffu) | (((uint32_t)im & 0xffu) << 8);
}

// Per-thread view of the ring block: a loader thread owns one antenna and
// fetches its dword (both pols, re/im) for a strided set of time rows.  The
// FPGA, delay, permutation and channel base are resolved once per thread.
// fetch() only issues loads (sample and scale dwords, no data-dependent
// branches, so all of a chunk's loads are in flight together); apply() scales
// them when the chunk is written to LDS.
template <bool PERM>
struct ThreadLoader {
  const uint32_t *chan32;  // channel base of the samples, dwords
  const uint16_t *chan16;  // same, 16-bit units (permutation path)
  const int16_t *cscale;   // channel base of the scales
  int a;                   // output antenna
  int f, n, off;           // identity path: FPGA, receiver in packet, sample offset
  int src[2], sp[2], pf[2], pn[2], poff[2];  // permutation path, per output pol

  __device__ void init(const VoltParams &P, int c, int a_) {
    a = a_;
    chan32 = reinterpret_cast<const uint32_t *>(P.samples) + (long)c * (P.NP + 2) * P.F * P.NT * P.RPP;
    chan16 = reinterpret_cast<const uint16_t *>(chan32);
    cscale = P.scales + (long)c * (P.NP + 2) * P.R * 2;
    f = a / P.RPP;
    n = a - f * P.RPP;
    off = P.window_base + (P.delays ? P.delays[f] : 0);
    if (PERM) {
#pragma unroll
      for (int po = 0; po < 2; ++po) {
        src[po] = P.recv_perm[a * 2 + po];
        sp[po] = P.pol_perm ? P.pol_perm[a * 2 + po] : po;
        const int s = src[po] < 0 ? 0 : src[po];
        pf[po] = s / P.RPP;
        pn[po] = s - pf[po] * P.RPP;
        poff[po] = P.window_base + (P.delays ? P.delays[pf[po]] : 0);
      }
    }
  }

  __device__ __forceinline__ void slot_of(const VoltParams &P, int g, int &slot, int &tt) const {
    if (P.nt_log2 >= 0) {
      slot = g >> P.nt_log2;
      tt = g & (P.NT - 1);
    } else {
      slot = g / P.NT;
      tt = g - slot * P.NT;
    }
  }

  // Issues the loads of window sample s: x = sample dword, sc = scale dword
  // (identity path) or x = two 16-bit samples, sc = two int16 scales (permutation).
  __device__ __forceinline__ void fetch(const VoltParams &P, int s, uint32_t &x, uint32_t &sc) const {
    if (!PERM) {
      int slot, tt;
      slot_of(P, s + off, slot, tt);
      x = chan32[(uint32_t)((slot * P.F + f) * P.NT + tt) * P.RPP + n];
      sc = reinterpret_cast<const uint32_t *>(cscale)[slot * P.R + a];
      return;
    }
    x = 0;
    sc = 0;
#pragma unroll
    for (int po = 0; po < 2; ++po) {
      const int ss = src[po] < 0 ? 0 : src[po];
      int slot, tt;
      slot_of(P, s + poff[po], slot, tt);
      const uint32_t hidx = ((uint32_t)((slot * P.F + pf[po]) * P.NT + tt) * P.RPP + pn[po]) * 2 + sp[po];
      const uint32_t v = chan16[hidx];
      const uint32_t s16 = (uint16_t)cscale[(slot * P.R + ss) * 2 + sp[po]];
      x |= (src[po] < 0 ? 0u : v) << (16 * po);
      sc |= (src[po] < 0 ? 0u : s16) << (16 * po);
    }
  }

  __device__ __forceinline__ static uint32_t apply(uint32_t x, uint32_t sc, bool &bad) {
    if (sc == 0x00010001u) return x;
    const int s0 = (int)(int16_t)(sc & 0xffffu), s1 = (int)(int16_t)(sc >> 16);
    return scale_pair(x & 0xffffu, s0, bad) | (scale_pair(x >> 16, s1, bad) << 16);
  }
};

template <int NG, bool DO_CORR, bool DO_BEAM, int CH>
struct VoltTraits {
  static constexpr int NA = 16 * NG;                       // padded antennas
  static constexpr int RSD = NA + 1;                       // LDS row stride in dwords (odd)
  static constexpr int NPAIRS = NG * (NG + 1) / 2;
  static constexpr int NWC = DO_CORR ? NPAIRS : 0;         // correlator waves, one per pair
  static constexpr int NWB = DO_BEAM ? (DO_CORR ? 2 : 4) : 0;  // beam waves: pol x time-tile parity
  static constexpr int NW = NWC + NWB;
  static constexpr int NT = 64 * NW;
};

// Loads component operands of antenna group G for this lane: op[v] = 16 bytes
// of component v (2*pol + re/im) of antenna 16G + (lane & 15) at times
// t0 + 16*(lane >> 4) + 0..15 of the chunk.
template <int RSD>
__device__ __forceinline__ void load_group(const uint32_t *lds, int G, int t0, int lane, v4i op[4]) {
  const int a = 16 * G + (lane & 15);
  const int tb = t0 + 16 * (lane >> 4);
  uint32_t x[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) x[e] = lds[(tb + e) * RSD + a];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    uint32_t y0, y1, y2, y3;
    transpose4x4(x[4 * j], x[4 * j + 1], x[4 * j + 2], x[4 * j + 3], y0, y1, y2, y3);
    op[0][j] = (int)y0;
    op[1][j] = (int)y1;
    op[2][j] = (int)y2;
    op[3][j] = (int)y3;
  }
}

__device__ __forceinline__ v4i mfma_i8(v4i a, v4i b, v4i c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// Pair index -> antenna groups (A, B), A <= B, row-major over the upper triangle.
template <int NG>
__device__ __forceinline__ void pair_groups(int pi, int &A, int &B) {
  A = 0;
  int rem = pi;
  while (rem >= NG - A) {
    rem -= NG - A;
    ++A;
  }
  B = A + rem;
}

template <int NG, int KODD, bool DO_CORR, bool DO_BEAM, int NMT, int CH, bool PERM>
__global__ void __launch_bounds__((VoltTraits<NG, DO_CORR, DO_BEAM, CH>::NT))
volt_kernel(VoltParams P) {
  using TR = VoltTraits<NG, DO_CORR, DO_BEAM, CH>;
  constexpr int RSD = TR::RSD;
  constexpr int NK = 2 * NG - KODD;  // beam K steps of 8 antennas
  __shared__ uint32_t lds[2][CH * RSD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int c, split;
  {
    const int bid = blockIdx.x;
    if (P.xcd_map) {
      const int xcd = bid & 7, k = bid >> 3;
      c = (k / P.nsplit) * 8 + xcd;
      split = k % P.nsplit;
    } else {
      c = bid / P.nsplit;
      split = bid % P.nsplit;
    }
  }
  const int s_begin = split * P.samples_per_wg;
  const int s_end = min(P.T, s_begin + P.samples_per_wg);
  if (s_begin >= s_end) return;
  const int nchunks = (s_end - s_begin + CH - 1) / CH;
  bool bad = false;
  const bool is_corr = wave < TR::NWC;
  const int bw = wave - TR::NWC;  // beam wave index 0..3 when >= 0

  // correlator accumulators of this wave's pair: [p][q][re, x, y]
  v4i acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = v4i{0, 0, 0, 0};
  int gA = 0, gB = 0;
  if (DO_CORR && is_corr) pair_groups<NG>(wave, gA, gB);

  // beam weight operands of this wave's pol, held for the whole workgroup
  const int bpol = bw & 1;
  v8h wop[DO_BEAM ? NMT : 1][DO_BEAM ? 2 * NK : 1];
  if constexpr (DO_BEAM) {
    if (!is_corr) {
      const v8h *wsrc = P.wop + (((long)c * 2 + bpol) * NMT) * NK * 2 * 64;
#pragma unroll
      for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
        for (int j = 0; j < NK; ++j) {
          wop[mt][2 * j] = wsrc[((mt * NK + j) * 2 + 0) * 64 + lane];
          wop[mt][2 * j + 1] = wsrc[((mt * NK + j) * 2 + 1) * 64 + lane];
        }
    }
  }

  // loader geometry: thread -> (antenna a = tid % R, row tr = tid / R)
  const int LT = TR::NT / P.R;  // time rows per pass (R <= 64 <= NT)
  const bool loader = tid < LT * P.R;
  ThreadLoader<PERM> ld;
  const int tr = tid / P.R;
  ld.init(P, c, tid - tr * P.R);
  constexpr int LTMIN = TR::NT / TR::NA;  // LT >= LTMIN since R <= NA
  constexpr int MAXLOAD = (CH + LTMIN - 1) / LTMIN;
  const int nload = (CH + LT - 1) / LT;
  uint32_t pre[MAXLOAD], psc[MAXLOAD];
  auto fetch = [&](int ch) {
    const int s0 = s_begin + ch * CH + tr;
#pragma unroll
    for (int k = 0; k < MAXLOAD; ++k) {
      if (k < nload) {
        const int t = tr + k * LT;
        const int s = s0 + k * LT;
        if (loader && t < CH && s < s_end) {
          ld.fetch(P, s, pre[k], psc[k]);
        } else {
          pre[k] = 0u;
          psc[k] = 0x00010001u;
        }
      }
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int k = 0; k < MAXLOAD; ++k) {
      if (k < nload) {
        const int t = tr + k * LT;
        if (loader && t < CH) lds[buf][t * RSD + ld.a] = ThreadLoader<PERM>::apply(pre[k], psc[k], bad);
      }
    }
  };
  // zero the padding antennas of both buffers once (the loader never writes them)
  for (int i = tid; i < 2 * CH * RSD; i += TR::NT) {
    const int a = (i % (CH * RSD)) % RSD;
    if (a >= P.R) (&lds[0][0])[i] = 0u;
  }

  fetch(0);
  stash(0);
  __syncthreads();

  for (int ch = 0; ch < nchunks; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nchunks) fetch(ch + 1);
    const uint32_t *L = lds[buf];

    if constexpr (DO_CORR) {
      if (is_corr) {
#pragma unroll 1
        for (int t0 = 0; t0 < CH; t0 += 64) {
          v4i opA[4], opB[4];
          load_group<RSD>(L, gA, t0, lane, opA);
          if (gB != gA) {
            load_group<RSD>(L, gB, t0, lane, opB);
          } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) opB[v] = opA[v];
          }
#pragma unroll
          for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
              v4i *a3 = &acc[(p * 2 + q) * 3];
              a3[0] = mfma_i8(opA[2 * p], opB[2 * q], a3[0]);
              a3[0] = mfma_i8(opA[2 * p + 1], opB[2 * q + 1], a3[0]);
              a3[1] = mfma_i8(opA[2 * p + 1], opB[2 * q], a3[1]);
              a3[2] = mfma_i8(opA[2 * p], opB[2 * q + 1], a3[2]);
            }
        }
      }
    }

    if constexpr (DO_BEAM) {
      if (!is_corr) {
        const int kg = lane >> 5;
#pragma unroll 1
        for (int tt = bw >> 1; tt < CH / 32; tt += TR::NWB / 2) {
          // time tile of 32 samples
          const int t = 32 * tt + (lane & 31);
          v16f bacc[NMT];
#pragma unroll
          for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
            for (int i = 0; i < 16; ++i) bacc[mt][i] = 0.0f;
#pragma unroll
          for (int j = 0; j < NK; ++j) {
            {
              const int a0 = 8 * j + 4 * kg;
              v8h xop;
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                const uint32_t pr = L[t * RSD + a0 + i] >> (16 * bpol);
                xop[2 * i] = (_Float16)(int)(int8_t)(pr & 0xffu);
                xop[2 * i + 1] = (_Float16)(int)(int8_t)((pr >> 8) & 0xffu);
              }
#pragma unroll
              for (int mt = 0; mt < NMT; ++mt) {
                bacc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(wop[mt][2 * j], xop, bacc[mt], 0, 0, 0);
                bacc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(wop[mt][2 * j + 1], xop, bacc[mt], 0, 0, 0);
              }
            }
          }
          const int s = s_begin + ch * CH + t;
          if (s < s_end) {
#pragma unroll
            for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
              for (int v = 0; v < 16; v += 2) {
                const int row = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int e = 16 * mt + (row >> 1);
                if (e < P.nbeams_out) {
                  const __half2 h = __floats2half2_rn(bacc[mt][v], bacc[mt][v + 1]);
                  reinterpret_cast<__half2 *>(P.beams)
                      [(((long)c * 2 + bpol) * P.nbeams_out + e) * P.T + s] = h;
                }
              }
          }
        }
      }
    }

    if (ch + 1 < nchunks) stash(buf ^ 1);
    __syncthreads();
  }

  if (bad) atomicOr(P.err, kErrRange);

  if constexpr (DO_CORR) {
    if (is_corr) {
      const long nbl = (long)P.R * (P.R + 1) / 2;
      const int b = 16 * gB + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int a = 16 * gA + 4 * (lane >> 4) + v;
        if (a > b || b >= P.R) continue;
        const long bl = (long)b * (b + 1) / 2 + a;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const v4i *a3 = &acc[(p * 2 + q) * 3];
            const int re = a3[0][v];
            const int im = a3[1][v] - a3[2][v];
            const long o = (((long)c * nbl + bl) * 4 + p * 2 + q) * 2;
            if (P.vis32) {
              atomicAdd(&P.vis32[o], re);
              atomicAdd(&P.vis32[o + 1], im);
            }
            if (P.vis64) {
              atomicAdd(reinterpret_cast<unsigned long long *>(&P.vis64[o]),
                        (unsigned long long)(long long)re);
              atomicAdd(reinterpret_cast<unsigned long long *>(&P.vis64[o + 1]),
                        (unsigned long long)(long long)im);
            }
          }
      }
    }
  }
}

template <int NG, int KODD, bool C_, bool B_, int NMT>
static hipError_t launch_t(const VoltParams &p, hipStream_t s) {
  constexpr int CH = 128;
  using TR = VoltTraits<NG, C_, B_, CH>;
  dim3 grid(p.C * p.nsplit), block(TR::NT);
  if (p.recv_perm)
    hipLaunchKernelGGL((volt_kernel<NG, KODD, C_, B_, NMT, CH, true>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((volt_kernel<NG, KODD, C_, B_, NMT, CH, false>), grid, block, 0, s, p);
  return hipGetLastError();
}

template <int NG, int KODD>
static hipError_t launch_ng(const VoltParams &p, bool do_corr, bool do_beam, hipStream_t s) {
  if (do_corr && do_beam) {
    if (p.NMT == 1) return launch_t<NG, KODD, true, true, 1>(p, s);
    if (p.NMT == 2) return launch_t<NG, KODD, true, true, 2>(p, s);
  } else if (do_corr) {
    return launch_t<NG, 0, true, false, 1>(p, s);
  } else if (do_beam) {
    if (p.NMT == 1) return launch_t<NG, KODD, false, true, 1>(p, s);
    if (p.NMT == 2) return launch_t<NG, KODD, false, true, 2>(p, s);
    if (p.NMT == 4) return launch_t<NG, KODD, false, true, 4>(p, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_volt(const VoltParams &p, bool do_corr, bool do_beam, hipStream_t s) {
  const int ng = (p.R + 15) / 16;
  const int nk = (p.R + 7) / 8;
  const int kodd = 2 * ng - nk;  // 0 or 1
  if (do_beam && p.NK != nk) return hipErrorInvalidValue;
#define LAM_NG(G)                                                   \
  case G:                                                           \
    return kodd ? launch_ng<G, 1>(p, do_corr, do_beam, s)           \
                : launch_ng<G, 0>(p, do_corr, do_beam, s);
  switch (ng) {
    LAM_NG(1)
    LAM_NG(2)
    LAM_NG(3)
    LAM_NG(4)
    default:
      return hipErrorInvalidValue;
  }
#undef LAM_NG
}

}  // namespace lam
