// Voltage kernels: ingest -> int8-MFMA correlator -> FP16-MFMA beamformer in
// one pass over the int8 voltages.
//
// Replaces, for one block, the reference chain
//   scale_and_convert_to_half (spatial.cuh:420-486) -> packet_to_corr_input
//   (spatial.cuh:215-304) -> TCC correlator (lambda_corr_beam_only_gpu_pipeline.hpp:314-316)
//   -> accumulate_visibilities_from_corr (src/spatial.cu:146-170)
//   -> packet_to_col_maj_cons (spatial.cuh:306-374) -> ccglib beam GEMM
//   (lambda_corr_beam_only_gpu_pipeline.hpp:335-338) -> beam_ccglib_to_half_output
//   (spatial.cuh:376-418)
// and the adaptive pipeline's apply_delays / reorder_streams
// (spatial.cuh:1420-1506, 593-649).
//
// Workgroup = (channel, time split).  128-sample chunks are staged in LDS as
// rows of [antenna] dwords (one dword = 2 pols x re/im int8) with a row stride
// of NA + 4 dwords (NA = antennas padded to a multiple of 16); that stride
// keeps both MFMA operand reads free of bank conflicts.  Padding columns hold
// finite garbage that only reaches discarded outputs or zero weights.
//
// Correlator: real rows are (antenna, component v = 2*pol + re/im).  Antennas
// form groups of 16; one wave per group pair (A <= B) runs
// v_mfma_i32_16x16x64_i8 on operands built with a 4x4 byte transpose.  Per
// (p, q) three int32 accumulators hold re = sum Re_a Re_b + Im_a Im_b,
// x = sum Im_a Re_b, y = sum Re_a Im_b; V = re + i (x - y) = sum x_a conj(x_b),
// exact, without negating operands.  Each workgroup stores its partial sums;
// a later pass adds the time splits.
//
// Beamformer: per (pol, 16-beam tile), [Re;Im] rows x K = (antenna, re/im) with
// v_mfma_f32_32x32x16_f16; weights arrive as hi + lo FP16 operands.
//
// Two loaders:
//   * fast (F == 1, R % 4 == 0, no permutation): a dedicated loader wave moves
//     the contiguous ring bytes straight into a 3-stage LDS ring with
//     global_load_lds_dwordx4 (no registers held, two chunks in flight);
//     non-unit scales and partial chunks are fixed up in LDS.
//   * general: every thread owns one antenna and prefetches its dwords into
//     registers (FPGA segments, per-FPGA delays, permutation, scales).
#include "lam_internal.hpp"

namespace lam {

// ------------------------------------------------------------ small helpers
__device__ __forceinline__ void transpose4x4(uint32_t x0, uint32_t x1, uint32_t x2,
                                             uint32_t x3, uint32_t &y0, uint32_t &y1,
                                             uint32_t &y2, uint32_t &y3) {
  // x[k] holds bytes (c0,c1,c2,c3) of time k; y[c] = bytes (t0,t1,t2,t3) of
  // component c.  v_perm_b32(hi, lo, sel) picks bytes of the 64-bit {hi, lo}.
  const uint32_t a_lo = __builtin_amdgcn_perm(x1, x0, 0x05010400u);
  const uint32_t a_hi = __builtin_amdgcn_perm(x1, x0, 0x07030602u);
  const uint32_t b_lo = __builtin_amdgcn_perm(x3, x2, 0x05010400u);
  const uint32_t b_hi = __builtin_amdgcn_perm(x3, x2, 0x07030602u);
  y0 = __builtin_amdgcn_perm(b_lo, a_lo, 0x05040100u);
  y1 = __builtin_amdgcn_perm(b_lo, a_lo, 0x07060302u);
  y2 = __builtin_amdgcn_perm(b_hi, a_hi, 0x05040100u);
  y3 = __builtin_amdgcn_perm(b_hi, a_hi, 0x07060302u);
}

// int8 (re, im) of one pol times an int16 scale; flags results outside [-128, 127].
__device__ __forceinline__ uint32_t scale_pair(uint32_t pair, int s, bool &bad) {
  const int re = (int)(int8_t)(pair & 0xffu) * s;
  const int im = (int)(int8_t)((pair >> 8) & 0xffu) * s;
  bad |= (re < -128) | (re > 127) | (im < -128) | (im > 127);
  return ((uint32_t)re & 0xffu) | (((uint32_t)im & 0xffu) << 8);
}

__device__ __forceinline__ uint32_t apply_scale(uint32_t x, uint32_t sc, bool &bad) {
  if (sc == 0x00010001u) return x;
  const int s0 = (int)(int16_t)(sc & 0xffffu), s1 = (int)(int16_t)(sc >> 16);
  return scale_pair(x & 0xffffu, s0, bad) | (scale_pair(x >> 16, s1, bad) << 16);
}

__device__ __forceinline__ void slot_of(const VoltParams &P, int g, int &slot, int &tt) {
  if (P.nt_log2 >= 0) {
    slot = g >> P.nt_log2;
    tt = g & (P.NT - 1);
  } else {
    slot = g / P.NT;
    tt = g - slot * P.NT;
  }
}

// s_waitcnt with only vmcnt constrained (gfx9 encoding: vmcnt[3:0] | vmcnt[5:4] << 14).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Workgroup barrier for the chunk loop, without the release fence of
// __syncthreads: that fence makes every wave drain vmcnt, which would stall
// the loader wave on the chunk it just prefetched and the beam waves on their
// global stores.  LDS hazards are covered explicitly: the loader waits for its
// LDS-DMA chunk with s_waitcnt vmcnt before arriving, and every wave drains
// lgkmcnt (its own LDS reads and writes) here.  The empty asm statements keep
// the compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void chunk_barrier() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt unconstrained
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// ------------------------------------------------------------- geometry
template <int NG, int KODD, bool DO_CORR, bool DO_BEAM, int NMT, bool LOADER_WAVE, bool TIGHT = false>
struct Geo {
  static constexpr int NA = 16 * NG;              // padded antennas
  // LDS row stride (dwords), == 4 mod 16.  TIGHT: R = NA - 12 antennas, rows
  // packed exactly as in the ring (reads of the padding columns run into the
  // next row: finite samples that only meet discarded outputs or zero weights)
  static constexpr int RSD = TIGHT ? NA - 12 : NA + 4;
  static constexpr int NK = 2 * NG - KODD;        // beam K steps of 8 antennas
  static constexpr int NPAIRS = NG * (NG + 1) / 2;
  static constexpr int NWC = DO_CORR ? NPAIRS : 0;        // correlator waves
  static constexpr int NWB = DO_BEAM ? 2 * NMT : 0;       // beam waves: pol x 16-beam tile
  static constexpr int NW = NWC + NWB + (LOADER_WAVE ? 1 : 0);
  static constexpr int NT = 64 * NW;
  static constexpr int STAGE = kChunk * RSD;      // dwords per staged chunk
  static_assert(STAGE % 256 == 0, "stage must be whole dwordx4 wave loads");
};

// ------------------------------------------------------------ correlator
// Operands of antenna group G: op[v] = 16 bytes of component v of antenna
// 16G + (lane & 15) at 16 times of the 64-row k-step t0.  Lane group g = lane >> 4
// reads rows t0 + 16*(e >> 2) + (e & 3) + 4g: with RSD == 4 (mod 16) the four
// groups hit disjoint banks.
template <int RSD>
__device__ __forceinline__ void load_group(const uint32_t *L, int G, int t0, int lane, v4i op[4]) {
  const int a = 16 * G + (lane & 15);
  const int tb = t0 + 4 * (lane >> 4);
  uint32_t x[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) x[e] = L[(tb + 16 * (e >> 2) + (e & 3)) * RSD + a];
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

// One chunk of correlation for the wave's group pair.
template <int RSD>
__device__ __forceinline__ void corr_chunk(const uint32_t *L, int gA, int gB, int lane, v4i acc[12]) {
#pragma unroll 1
  for (int t0 = 0; t0 < kChunk; t0 += 64) {
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

// Operand tiles of one chunk, transposed once per workgroup: tile
// (G, ks, v) holds op[v] of antenna group G at k-step ks (rows 64 ks ..) for
// every lane, [group][k-step][component][lane].  Wave u < 2 NG builds tile
// set u = 2 G + ks (the correlator waves, and the loader wave when NG = 1),
// so each group is transposed once per chunk instead of once per group pair.
template <int RSD, int NG>
__device__ __forceinline__ void transpose_chunk(const uint32_t *L, v4i *tiles, int wave, int lane) {
  if (wave < 2 * NG) {
    v4i op[4];
    load_group<RSD>(L, wave >> 1, 64 * (wave & 1), lane, op);
#pragma unroll
    for (int v = 0; v < 4; ++v) tiles[(wave * 4 + v) * 64 + lane] = op[v];
  }
}

// One chunk of correlation for the wave's group pair, from the transposed tiles.
__device__ __forceinline__ void corr_chunk_tiles(const v4i *tiles, int gA, int gB, int lane, v4i acc[12]) {
  static_assert(kChunk == 128, "two 64-row k-steps per chunk");
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    v4i opA[4], opB[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      opA[v] = tiles[((2 * gA + ks) * 4 + v) * 64 + lane];
      opB[v] = tiles[((2 * gB + ks) * 4 + v) * 64 + lane];
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

// Partial visibilities of this workgroup: vis_part[split][c][p*2+q][bl(a,b)][re,im]
// (polarisation pair major, so the XX and YY covariances are contiguous runs).
__device__ __forceinline__ void corr_store(const VoltParams &P, const v4i acc[12], int gA, int gB,
                                           int lane, int c, int split) {
  const long nbl = (long)P.R * (P.R + 1) / 2;
  int *base = P.vis_part + ((long)split * P.C + c) * nbl * 8;
  const int b = 16 * gB + (lane & 15);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int a = 16 * gA + 4 * (lane >> 4) + v;
    if (a > b || b >= P.R) continue;
    int2 *o = reinterpret_cast<int2 *>(base) + (long)b * (b + 1) / 2 + a;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const v4i *a3 = &acc[(p * 2 + q) * 3];
        o[(p * 2 + q) * nbl] = make_int2(a3[0][v], a3[1][v] - a3[2][v]);
      }
  }
}

// ------------------------------------------------------------ beamformer
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

// N (<= 4) dwordx4 beam operands at LDS byte address a, a + 32, ...: the
// dwords of 4 antennas x both pols for consecutive K steps.  The reads and
// their wait sit in one asm block.  Read through C++, the compiler would drain
// vmcnt before them (they may alias the loader wave's LDS-DMA writes), and for
// a beam wave vmcnt counts its own pending beam stores; the loader's vmcnt
// wait plus the chunk barrier already order the DMA before these reads.
template <int N>
__device__ __forceinline__ void lds_read_ops(uint32_t a, v4u *d) {
  static_assert(N >= 1 && N <= 4, "group of K steps");
  if constexpr (N == 1) {
    asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=&v"(d[0]) : "v"(a) : "memory");
  } else if constexpr (N == 2) {
    asm volatile("ds_read_b128 %0, %2\n\tds_read_b128 %1, %2 offset:32\n\ts_waitcnt lgkmcnt(0)"
                 : "=&v"(d[0]), "=&v"(d[1]) : "v"(a) : "memory");
  } else if constexpr (N == 3) {
    asm volatile(
        "ds_read_b128 %0, %3\n\tds_read_b128 %1, %3 offset:32\n\tds_read_b128 %2, %3 offset:64\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(d[0]), "=&v"(d[1]), "=&v"(d[2]) : "v"(a) : "memory");
  } else {
    asm volatile(
        "ds_read_b128 %0, %4\n\tds_read_b128 %1, %4 offset:32\n\tds_read_b128 %2, %4 offset:64\n\t"
        "ds_read_b128 %3, %4 offset:96\n\ts_waitcnt lgkmcnt(0)"
        : "=&v"(d[0]), "=&v"(d[1]), "=&v"(d[2]), "=&v"(d[3]) : "v"(a) : "memory");
  }
}

__device__ __forceinline__ uint32_t lds_addr(const uint32_t *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) uint32_t *)(p);
}

// (re, im) int8 of one pol (byte pair `sel` picks) as exact FP16: the bytes,
// offset to unsigned, become the low mantissa bits of 1024 (f16 0x6400), and
// one packed subtraction of 1152 restores the signed values.
__device__ __forceinline__ h2v i8pair_to_f16(uint32_t w, uint32_t sel) {
  const uint32_t m = __builtin_amdgcn_perm(0x64646464u, w ^ 0x80808080u, sel);
  return __builtin_bit_cast(h2v, m) - h2v{(_Float16)1152.0f, (_Float16)1152.0f};
}

// B operand of one K step: (re, im) int8 of 4 antennas of pol `pol`, as
// FP16 (exact) or BF16 (exact) bits.
template <int WP>
__device__ __forceinline__ v8h beam_operand(v4u d, int pol, uint32_t sel) {
  if constexpr (WP == kWprecBF16) {
    v8bf xb;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t pr = d[i] >> (16 * pol);
      xb[2 * i] = (__bf16)(float)(int)(int8_t)(pr & 0xffu);
      xb[2 * i + 1] = (__bf16)(float)(int)(int8_t)((pr >> 8) & 0xffu);
    }
    return __builtin_bit_cast(v8h, xb);
  } else {
    v8h x;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const h2v h = i8pair_to_f16(d[i], sel);
      x[2 * i] = h[0];
      x[2 * i + 1] = h[1];
    }
    return x;
  }
}

template <int WP>
__device__ __forceinline__ v16f beam_mfma(v8h w, v8h x, v16f acc) {
  if constexpr (WP == kWprecBF16)
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, w), __builtin_bit_cast(v8bf, x), acc,
                                                   0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(w, x, acc, 0, 0, 0);
}

// One chunk of beams for (pol, 16-beam tile mt): kChunk/32 time tiles of 32
// samples, TT tiles at a time (independent accumulators keep the matrix core
// busy while the next operands convert).  GK: K steps whose operands are read
// together in one asm block (fewer when the kernel also holds correlator
// state, to stay within the register budget); GK == 0: plain LDS reads
// (kernels without an LDS-DMA loader).  WP: weight precision.
template <int RSD, int NK, int GK, int TT, int WP>
__device__ __forceinline__ void beam_tiles(const VoltParams &P, const uint32_t *L, const v8h wop[2 * NK],
                                           int lane, int c, int pol, int mt, int s_chunk, int s_end) {
  static_assert((kChunk / 32) % TT == 0, "tiles per pass");
  constexpr int G = GK < 1 ? 1 : (GK < 4 ? GK : 4);
  const int kg = lane >> 5;
  const uint32_t sel = pol ? 0x04030402u : 0x04010400u;
  const bool full_tile = 16 * (mt + 1) <= P.nbeams_out;
  __half2 *const prow = reinterpret_cast<__half2 *>(P.beams) + ((long)c * 2 + pol) * P.nbeams_out * P.T;
#pragma unroll 1
  for (int t0 = 0; t0 < kChunk / 32; t0 += TT) {
    v4u d[TT][NK];
#pragma unroll
    for (int u = 0; u < TT; ++u) {
      const int t = 32 * (t0 + u) + (lane & 31);
      if constexpr (GK == 0) {
#pragma unroll
        for (int j = 0; j < NK; ++j) d[u][j] = *reinterpret_cast<const v4u *>(L + t * RSD + 8 * j + 4 * kg);
      } else {
        const uint32_t a = lds_addr(L + t * RSD + 4 * kg);
#pragma unroll
        for (int j = 0; j < NK; j += G) {
          if (NK - j >= G) lds_read_ops<G>(a + 32 * j, &d[u][j]);
          else if (NK - j == 3) lds_read_ops<(G >= 3 ? 3 : 1)>(a + 32 * j, &d[u][j]);
          else if (NK - j == 2) lds_read_ops<(G >= 2 ? 2 : 1)>(a + 32 * j, &d[u][j]);
          else lds_read_ops<1>(a + 32 * j, &d[u][j]);
        }
      }
    }
    v16f acc[TT];
#pragma unroll
    for (int u = 0; u < TT; ++u)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[u][i] = 0.0f;
#pragma unroll
    for (int j = 0; j < NK; ++j) {
      v8h x[TT];
#pragma unroll
      for (int u = 0; u < TT; ++u) x[u] = beam_operand<WP>(d[u][j], pol, sel);
#pragma unroll
      for (int u = 0; u < TT; ++u) acc[u] = beam_mfma<WP>(wop[2 * j], x[u], acc[u]);
      if constexpr (WP == kWprecF16HiLo) {
#pragma unroll
        for (int u = 0; u < TT; ++u) acc[u] = beam_mfma<WP>(wop[2 * j + 1], x[u], acc[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < TT; ++u) {
      const int s = s_chunk + 32 * (t0 + u) + (lane & 31);
      if (s < s_end) {
        __half2 *row = prow + s;
        if (full_tile) {
#pragma unroll
          for (int v = 0; v < 16; v += 2) {
            const int r = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            row[(long)(16 * mt + (r >> 1)) * P.T] = __floats2half2_rn(acc[u][v], acc[u][v + 1]);
          }
        } else {
#pragma unroll
          for (int v = 0; v < 16; v += 2) {
            const int r = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            const int e = 16 * mt + (r >> 1);
            if (e < P.nbeams_out) row[(long)e * P.T] = __floats2half2_rn(acc[u][v], acc[u][v + 1]);
          }
        }
      }
    }
  }
}

template <int RSD, int NK, int GK, int TT>
__device__ __forceinline__ void beam_chunk(const VoltParams &P, const uint32_t *L, const v8h wop[2 * NK],
                                           int lane, int c, int pol, int mt, int s_chunk, int s_end) {
  if (P.wprec == kWprecF16HiLo)
    beam_tiles<RSD, NK, GK, TT, kWprecF16HiLo>(P, L, wop, lane, c, pol, mt, s_chunk, s_end);
  else if (P.wprec == kWprecF16)
    beam_tiles<RSD, NK, GK, TT, kWprecF16>(P, L, wop, lane, c, pol, mt, s_chunk, s_end);
  else
    beam_tiles<RSD, NK, GK, TT, kWprecBF16>(P, L, wop, lane, c, pol, mt, s_chunk, s_end);
}

template <int NK>
__device__ __forceinline__ void load_weights(const VoltParams &P, int c, int pol, int mt, int NMT, int lane,
                                             v8h wop[2 * NK]) {
  const v8h *src = P.wop + ((((long)c * 2 + pol) * NMT + mt) * NK) * 2 * 64;
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    wop[2 * j] = src[(j * 2 + 0) * 64 + lane];
    wop[2 * j + 1] = src[(j * 2 + 1) * 64 + lane];
  }
}

__device__ __forceinline__ void wg_coords(const VoltParams &P, int &c, int &split) {
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

// ================================================================ fast path
#ifndef LAM_VOLT_STAGES
#define LAM_VOLT_STAGES 3
#endif
// ring depth of correlator-only workgroups (every stage holds a chunk in flight):
// two stages (61.5 KB of LDS with the operand tiles, against 80 KB for three) ran the
// correlator 147 against 154 us (384 ch x 8192 samples, one split) with the
// overlapped step the same or ~1 % shorter
#ifndef LAM_CORR_STAGES
#define LAM_CORR_STAGES 2
#endif
template <int NG, int KODD, bool DO_CORR, bool DO_BEAM, int NMT, bool TIGHT>
__global__ void __launch_bounds__((Geo<NG, KODD, DO_CORR, DO_BEAM, NMT, true, TIGHT>::NT))
volt_fast_kernel(VoltParams P) {
  using G = Geo<NG, KODD, DO_CORR, DO_BEAM, NMT, true, TIGHT>;
  constexpr int RSD = G::RSD, NK = G::NK, STAGE = G::STAGE;
  constexpr int NI = STAGE / 256;  // dwordx4 wave loads per chunk
  // LDS ring depth: up to LAM_VOLT_STAGES, as long as the wait count stays encodable
  constexpr int NS = (DO_CORR && !DO_BEAM) ? LAM_CORR_STAGES : ((LAM_VOLT_STAGES >= 4 && 2 * NI < 64) ? 4 : 3);
  static_assert(NS >= 2 && NS <= 4, "ring depth");
  static_assert(NI < 64, "loads of one chunk must fit the vmcnt counter");
  static_assert(NI * (NS - 2) < 64, "the wait count for the oldest chunk must fit vmcnt");
  // Chunks the loader keeps in flight.  A correlator-only workgroup is done with a raw
  // stage once the stage is transposed into the operand tiles, so every stage can hold
  // a chunk in flight; the beam waves read the raw stage while they work on it, so
  // kernels with beams keep one stage for the chunk being processed.
  constexpr int PF = (DO_CORR && !DO_BEAM && NI * (NS - 1) < 64) ? NS : NS - 1;
  // slack past the last stage for the padding-column reads of a TIGHT layout
  __shared__ __align__(16) uint32_t lds[NS * STAGE + (TIGHT ? 16 : 0)];
  __shared__ v4i tiles[DO_CORR ? NG * 2 * 4 * 64 : 1];  // the chunk's correlator operands
  __shared__ int s_flags;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int c, split;
  wg_coords(P, c, split);
  const int s_begin = split * P.samples_per_wg;
  const int s_end = min(P.T, s_begin + P.samples_per_wg);
  if (s_begin >= s_end) return;
  const int nchunks = (s_end - s_begin + kChunk - 1) / kChunk;
  const bool is_loader = wave == G::NW - 1;
  const bool is_corr = wave < G::NWC;
  const int bw = wave - G::NWC;
  const int off = P.window_base + (P.delays ? P.delays[0] : 0);
  const long ring_rows = (long)(P.NP + 2) * P.NT;
  const uint32_t *chan32 = reinterpret_cast<const uint32_t *>(P.samples) + (long)c * ring_rows * P.R;
  const uint32_t *cscale = reinterpret_cast<const uint32_t *>(P.scales) + (long)c * (P.NP + 2) * P.R;
  bool bad = false;

  // any non-unit scale in this workgroup's slot range?
  if (tid == 0) s_flags = 0;
  __syncthreads();
  {
    const int slot0 = (s_begin + off) / P.NT, slot1 = (s_end - 1 + off) / P.NT;
    const int n = (slot1 - slot0 + 1) * P.R;
    int any = 0;
    for (int i = tid; i < n; i += G::NT) any |= cscale[(long)slot0 * P.R + i] != 0x00010001u;
    if (any) atomicOr(&s_flags, 1);
  }
  __syncthreads();
  const bool scaled = s_flags != 0;

  auto issue = [&](int ch) {
    const int s0 = s_begin + ch * kChunk;
    const int valid = min(kChunk, s_end - s0);
    uint32_t *dst = lds + (ch % NS) * STAGE;
    if (TIGHT && valid == kChunk) {
      // rows packed as in the ring: one contiguous run of STAGE dwords
      const uint32_t *base = chan32 + (long)(s0 + off) * P.R;
#pragma unroll
      for (int j = 0; j < NI; ++j)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(base + j * 256 + lane * 4),
                                         (__attribute__((address_space(3))) void *)(dst + j * 256), 16, 0, 0);
      return;
    }
#pragma unroll 2
    for (int j = 0; j < NI; ++j) {
      const int pos = j * 256 + lane * 4;
      const int t = pos / RSD, a = pos - t * RSD;
      const uint32_t *src = chan32;
      if (t < valid && a < P.R) src = chan32 + (long)(s0 + t + off) * P.R + a;
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src),
                                       (__attribute__((address_space(3))) void *)(dst + j * 256), 16, 0, 0);
    }
  };

  v4i acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = v4i{0, 0, 0, 0};
  int gA = 0, gB = 0;
  if (DO_CORR && is_corr) pair_groups<NG>(wave, gA, gB);
  const int bpol = bw & 1, bmt = bw >> 1;
  v8h wop[DO_BEAM ? 2 * NK : 1];
  if constexpr (DO_BEAM) {
    if (!is_corr && !is_loader) {
      load_weights<NK>(P, c, bpol, bmt, NMT, lane, wop);
      // the weights arrive before the chunk loop: inside it a beam wave's
      // vmcnt then counts only its own beam stores, which nothing waits for
      wait_vmcnt<0>();
    }
  }

  if (is_loader) {
#pragma unroll
    for (int k = 0; k < PF; ++k)
      if (k < nchunks) issue(k);
  }
  for (int ch = 0; ch < nchunks; ++ch) {
    if (is_loader) {
      // chunk ch is complete when at most min(PF - 1, chunks issued after it) chunk loads remain
      const int after = min(PF - 1, nchunks - 1 - ch);
      if constexpr (PF >= 4) {
        if (after >= 3) wait_vmcnt<(PF >= 4 ? 3 * NI : 0)>();
        else if (after == 2) wait_vmcnt<2 * NI>();
        else if (after == 1) wait_vmcnt<NI>();
        else wait_vmcnt<0>();
      } else if constexpr (PF == 3) {
        if (after >= 2) wait_vmcnt<2 * NI>();
        else if (after == 1) wait_vmcnt<NI>();
        else wait_vmcnt<0>();
      } else {
        if (after >= 1) wait_vmcnt<NI>();
        else wait_vmcnt<0>();
      }
    }
    uint32_t *L = lds + (ch % NS) * STAGE;
    const int s0 = s_begin + ch * kChunk;
    const int valid = min(kChunk, s_end - s0);
    if (scaled || (DO_CORR && valid < kChunk)) {
      __syncthreads();  // raw chunk visible to all waves
      for (int i = tid; i < kChunk * P.R; i += G::NT) {
        const int t = i / P.R, a = i - t * P.R;
        uint32_t &w = L[t * RSD + a];
        if (t >= valid) {
          w = 0u;
        } else if (scaled) {
          int slot, tt;
          slot_of(P, s0 + t + off, slot, tt);
          w = apply_scale(w, cscale[(long)slot * P.R + a], bad);
        }
      }
    }
    chunk_barrier();
    if constexpr (DO_CORR) {
      // (the previous chunk's tile reads ended before the barrier above)
      transpose_chunk<RSD, NG>(L, tiles, wave, lane);
      chunk_barrier();
    }
    // (the refilled stage's last readers, the transpose or the previous chunk's beam
    // waves, finished before the barrier above)
    if (is_loader && ch + PF < nchunks) issue(ch + PF);
    if constexpr (DO_CORR) {
      if (is_corr) corr_chunk_tiles(tiles, gA, gB, lane, acc);
    }
    if constexpr (DO_BEAM) {
      if (!is_corr && !is_loader)
        beam_chunk<RSD, NK, DO_CORR ? 1 : 4, (DO_CORR || NK > 5) ? 1 : 2>(P, L, wop, lane, c, bpol, bmt, s0, s_end);
    }
  }
  if (bad) atomicOr(P.err, kErrRange);
  if constexpr (DO_CORR) {
    if (is_corr) corr_store(P, acc, gA, gB, lane, c, split);
  }
}

// ============================================================= general path
// Per-thread view of the ring block: a loader thread owns one antenna and
// fetches its dword (both pols, re/im) for a strided set of time rows; FPGA,
// delay, permutation and channel base are resolved once per thread.
template <bool PERM>
struct ThreadLoader {
  const uint32_t *chan32;
  const uint16_t *chan16;
  const int16_t *cscale;
  int a, f, n, off;
  int src[2], sp[2], pf[2], pn[2], poff[2];

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

  // Issues the loads of window sample s (no data-dependent branches).
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
};

template <int NG, int KODD, bool DO_CORR, bool DO_BEAM, int NMT, bool PERM>
__global__ void __launch_bounds__((Geo<NG, KODD, DO_CORR, DO_BEAM, NMT, false>::NT))
volt_general_kernel(VoltParams P) {
  using G = Geo<NG, KODD, DO_CORR, DO_BEAM, NMT, false>;
  constexpr int RSD = G::RSD, NK = G::NK, STAGE = G::STAGE;
  __shared__ __align__(16) uint32_t lds[2 * STAGE];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int c, split;
  wg_coords(P, c, split);
  const int s_begin = split * P.samples_per_wg;
  const int s_end = min(P.T, s_begin + P.samples_per_wg);
  if (s_begin >= s_end) return;
  const int nchunks = (s_end - s_begin + kChunk - 1) / kChunk;
  bool bad = false;
  const bool is_corr = wave < G::NWC;
  const int bw = wave - G::NWC;

  v4i acc[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) acc[i] = v4i{0, 0, 0, 0};
  int gA = 0, gB = 0;
  if (DO_CORR && is_corr) pair_groups<NG>(wave, gA, gB);
  const int bpol = bw & 1, bmt = bw >> 1;
  v8h wop[DO_BEAM ? 2 * NK : 1];
  if constexpr (DO_BEAM) {
    if (!is_corr) load_weights<NK>(P, c, bpol, bmt, NMT, lane, wop);
  }

  // loader geometry: thread -> (antenna a = tid % R, row tr = tid / R)
  const int LT = G::NT / P.R;
  const bool loader = tid < LT * P.R;
  ThreadLoader<PERM> ld;
  const int tr = tid / P.R;
  ld.init(P, c, tid - tr * P.R);
  constexpr int LTMIN = G::NT / G::NA;
  constexpr int MAXLOAD = (kChunk + LTMIN - 1) / LTMIN;
  const int nload = (kChunk + LT - 1) / LT;
  uint32_t pre[MAXLOAD], psc[MAXLOAD];
  auto fetch = [&](int ch) {
    const int s0 = s_begin + ch * kChunk + tr;
#pragma unroll
    for (int k = 0; k < MAXLOAD; ++k) {
      if (k < nload) {
        const int t = tr + k * LT;
        const int s = s0 + k * LT;
        if (loader && t < kChunk && s < s_end) {
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
        if (loader && t < kChunk) lds[buf * STAGE + t * RSD + ld.a] = apply_scale(pre[k], psc[k], bad);
      }
    }
  };
  // padding antennas are never loaded: clear them once in both buffers
  for (int i = tid; i < 2 * STAGE; i += G::NT)
    if ((i % RSD) >= P.R) lds[i] = 0u;

  fetch(0);
  stash(0);
  __syncthreads();
  for (int ch = 0; ch < nchunks; ++ch) {
    const int buf = ch & 1;
    if (ch + 1 < nchunks) fetch(ch + 1);
    const uint32_t *L = lds + buf * STAGE;
    const int s0 = s_begin + ch * kChunk;
    if constexpr (DO_CORR) {
      if (is_corr) corr_chunk<RSD>(L, gA, gB, lane, acc);
    }
    if constexpr (DO_BEAM) {
      if (!is_corr) beam_chunk<RSD, NK, 0, 1>(P, L, wop, lane, c, bpol, bmt, s0, s_end);
    }
    if (ch + 1 < nchunks) stash(buf ^ 1);
    __syncthreads();
  }
  if (bad) atomicOr(P.err, kErrRange);
  if constexpr (DO_CORR) {
    if (is_corr) corr_store(P, acc, gA, gB, lane, c, split);
  }
}

// ================================================================ launchers
template <int NG, int KODD, bool C_, bool B_, int NMT>
static hipError_t launch_t(const VoltParams &p, hipStream_t s) {
  dim3 grid(p.C * p.nsplit);
  const bool fast = p.F == 1 && (p.R % 4) == 0 && p.recv_perm == nullptr;
  if (fast && p.R == 16 * NG - 12) {
    using G = Geo<NG, KODD, C_, B_, NMT, true, true>;
    hipLaunchKernelGGL((volt_fast_kernel<NG, KODD, C_, B_, NMT, true>), grid, dim3(G::NT), 0, s, p);
  } else if (fast) {
    using G = Geo<NG, KODD, C_, B_, NMT, true>;
    hipLaunchKernelGGL((volt_fast_kernel<NG, KODD, C_, B_, NMT, false>), grid, dim3(G::NT), 0, s, p);
  } else {
    using G = Geo<NG, KODD, C_, B_, NMT, false>;
    if (p.recv_perm)
      hipLaunchKernelGGL((volt_general_kernel<NG, KODD, C_, B_, NMT, true>), grid, dim3(G::NT), 0, s, p);
    else
      hipLaunchKernelGGL((volt_general_kernel<NG, KODD, C_, B_, NMT, false>), grid, dim3(G::NT), 0, s, p);
  }
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
  if (p.samples_per_wg % kChunk != 0) return hipErrorInvalidValue;
#define LAM_NG(G)                                         \
  case G:                                                 \
    return kodd ? launch_ng<G, 1>(p, do_corr, do_beam, s) \
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

// Sum of the per-split partial visibilities ([split][c][pq][bl][re,im]) into the
// int64 accumulator [c][bl][pq][re,im] (atomic when several buffers' blocks may
// be in flight on different streams).  i walks the accumulator, so its
// read-modify-write is contiguous; a wave reads four 64-byte runs of each
// split.  nbl8 = 8 * baselines.
template <bool ATOMIC>
__global__ void accumulate_vis_kernel(const int *part, int nsplit, long long count, int nbl8, long long *acc) {
  const int nbl = nbl8 >> 3;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    const long long c = i / nbl8;
    const int r = (int)(i - c * nbl8);
    const int bl = r >> 3, pq = (r >> 1) & 3, ri = r & 1;
    const long long src = c * nbl8 + ((long long)pq * nbl + bl) * 2 + ri;
    long long s = 0;
    for (int k = 0; k < nsplit; ++k) s += part[(long long)k * count + src];
    if (ATOMIC)
      atomicAdd(reinterpret_cast<unsigned long long *>(acc + i), (unsigned long long)s);
    else
      acc[i] += s;
  }
}

hipError_t launch_accumulate_vis(const int *vis_part, int nsplit, long long count, int nbl, long long *acc,
                                 bool atomic, hipStream_t s) {
  const int blocks = (int)std::min<long long>((count + 255) / 256, 8192);
  if (atomic)
    hipLaunchKernelGGL(accumulate_vis_kernel<true>, dim3(blocks), dim3(256), 0, s, vis_part, nsplit, count, 8 * nbl,
                       acc);
  else
    hipLaunchKernelGGL(accumulate_vis_kernel<false>, dim3(blocks), dim3(256), 0, s, vis_part, nsplit, count,
                       8 * nbl, acc);
  return hipGetLastError();
}

}  // namespace lam
