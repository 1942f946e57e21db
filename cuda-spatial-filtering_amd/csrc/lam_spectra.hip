// Post-beam spectra of the adaptive pipeline: per (channel, pol, beam) the
// 64-point forward FFT of every packet of beam samples, the five fine
// channels on each side of the coarse-channel edge removed, and the FFT
// magnitudes summed over the block's packets.
//
// Replaces the reference's beamToCUFFTInput permutation -> cufftXtExec
// (64-point C2C, batch 2B*C*P*NP) -> cufftOutputToFineChannelRemove -> two
// half copies -> fineChannelRemovedToBeamFFTDownsample -> sum_fft_over_packets
// (lambda_adaptive_beamformed_spectra_pipeline.hpp:808-848, 1046-1062;
// spatial.cuh:1508-1567).  Output bin g of NT - 10 = 54: fine channel
// u = g + 37 for g < 27 (negative frequencies -27..-1), u = g - 27 otherwise
// (0..26); u = 27..36 are dropped.
//
// The transform is a DFT on the matrix cores rather than a radix FFT: as a
// real GEMM it is X (re and im rows of 54 bins) = F (K = 128: 64 samples x
// re, im) x B (one packet per column), v_mfma_f32_16x16x32_f16 with 16 bins
// x 16 packets per tile.  The beams are FP16 already, so B is the FP16 beam
// output as stored; F is split into FP16 hi + lo parts (two MFMAs each), so
// the products carry ~22 bits of F and the FP32 accumulation gives FFT-grade
// accuracy.  Only the real rows of F are stored: for bin u,
//   F[u][2t] = cos(2 pi u t / 64),  F[u][2t + 1] = sin(2 pi u t / 64),
// X_re = F B, and X_im = F B' with B'[2t] = x_im, B'[2t+1] = -x_re (a half
// swap and a sign flip of each packed sample).  A radix FFT over lanes costs
// ~65 VALU instructions per packet and ran ALU-bound at under half the HBM
// rate; here the matrix cores need ~2/3 of the HBM time for the same data.
//
// Work: one wave per (channel, pol, beam) row at a time, persistent
// workgroups of four waves, F (32 KB, both parts) in LDS.  Packet tiles of 16
// columns, the next tile's loads in flight during the current tile's MFMAs;
// 4 bin tiles x (re, im) give eight independent accumulators.  Lane n of each
// quarter-wave accumulates |X| of its packet for the 16 bins it holds, and a
// reduce-scatter over the 16 lanes at the end of the row leaves one bin per
// lane.
#include <cmath>

#include "lam_internal.hpp"

namespace lam {

constexpr int kSpecFFT = 64, kSpecDrop = 5;
constexpr int kSpecKeep = kSpecFFT / 2 - kSpecDrop;  // 27 bins on each side of DC
constexpr int kSpecOut = kSpecFFT - 2 * kSpecDrop;   // 54
constexpr int kSpecWaves = 4, kSpecMT = 4, kSpecKS = 4;  // 4 tiles of 16 bins, 4 K steps of 32
constexpr int kSpecFrags = kSpecMT * kSpecKS * 2 * 64;  // 16-byte A fragments (hi, lo)
typedef unsigned su4 __attribute__((ext_vector_type(4)));
typedef float sf4 __attribute__((ext_vector_type(4)));

size_t spectra_table_bytes() { return (size_t)kSpecFrags * 16; }

// A-operand fragments of F, [mt][ks][hi/lo][lane][8 halves]: lane l holds row
// (bin tile row) l % 16 and K = 32 ks + 8 (l / 16) + j.  Output bin g = 16 mt
// + row maps to fine channel u = g + 37 (g < 27) or g - 27; rows g >= 54 are 0.
void spectra_table_host(uint16_t *dst) {
  for (int mt = 0; mt < kSpecMT; ++mt)
    for (int ks = 0; ks < kSpecKS; ++ks)
      for (int lane = 0; lane < 64; ++lane)
        for (int j = 0; j < 8; ++j) {
          const int g = 16 * mt + (lane & 15), k = 32 * ks + 8 * (lane >> 4) + j, t = k >> 1;
          double v = 0.0;
          if (g < kSpecOut) {
            const int u = g < kSpecKeep ? g + kSpecFFT - kSpecKeep : g - kSpecKeep;
            const double a = 2.0 * M_PI * (double)((u * t) % kSpecFFT) / kSpecFFT;
            v = (k & 1) ? std::sin(a) : std::cos(a);
          }
          const __half hi = __float2half((float)v);
          const __half lo = __float2half((float)(v - (double)__half2float(hi)));
          const long base = (((long)(mt * kSpecKS + ks) * 2) * 64 + lane) * 8 + j;
          dst[base] = __half_as_ushort(hi);
          dst[base + 64 * 8] = __half_as_ushort(lo);
        }
}

// Samples 16 s + 4 kq .. + 3 of packet pk of a row, s = 0..3 (the B operand of
// K step s for column pk).  Loads are unconditional (a partial tile's absent
// packets re-read the row's last packet) so that no branch splits the
// prefetch; those columns are zeroed when the tile is used.
__device__ __forceinline__ void spectra_load(const su4 *__restrict__ beams, long row, int np, int pk, int kq,
                                             su4 d[kSpecKS]) {
  const su4 *x = beams + (row * np + (pk < np ? pk : np - 1)) * 16;
#pragma unroll
  for (int s = 0; s < kSpecKS; ++s) d[s] = x[4 * s + kq];
}

__global__ void __launch_bounds__(64 * kSpecWaves, 4)
spectra_kernel(const su4 *__restrict__ beams, const su4 *__restrict__ table, int rows, int np,
               float *__restrict__ out) {
  __shared__ su4 F[kSpecFrags];
  for (int i = threadIdx.x; i < kSpecFrags; i += 64 * kSpecWaves) F[i] = table[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  const int ntile = (np + 15) / 16;
  const long stride = (long)gridDim.x * kSpecWaves;
  long row = (long)blockIdx.x * kSpecWaves + w;
  int tile = 0;
  float sum[kSpecMT][4];
#pragma unroll
  for (int mt = 0; mt < kSpecMT; ++mt)
#pragma unroll
    for (int i = 0; i < 4; ++i) sum[mt][i] = 0.0f;
  su4 cur[kSpecKS], nxt[kSpecKS];
  if (row < rows) spectra_load(beams, row, np, n, kq, cur);
  while (row < rows) {
    long nrow = row;
    int ntl = tile + 1;
    if (ntl == ntile) {
      nrow = row + stride;
      ntl = 0;
    }
    // the next tile (possibly the next row's first) is in flight during this
    // one's MFMAs; unconditional (the last tile re-reads its own row)
    spectra_load(beams, nrow < rows ? nrow : row, np, 16 * ntl + n, kq, nxt);
    if (16 * tile + n >= np) {  // absent packets of a partial tile add |0|
#pragma unroll
      for (int s = 0; s < kSpecKS; ++s) cur[s] = su4{0u, 0u, 0u, 0u};
    }
    // opaque per K step, so the fragment reads stay next to their MFMAs
    // instead of being hoisted out of the loop into registers
    int fl_lane = lane;
    sf4 are[kSpecMT], aim[kSpecMT];
#pragma unroll
    for (int mt = 0; mt < kSpecMT; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) are[mt][i] = aim[mt][i] = 0.0f;
#pragma unroll
    for (int s = 0; s < kSpecKS; ++s) {
      su4 dp;  // (re, im) -> (im, -re) per packed sample
#pragma unroll
      for (int q = 0; q < 4; ++q) dp[q] = __builtin_amdgcn_alignbit(cur[s][q], cur[s][q], 16) ^ 0x80000000u;
      const v8h b = __builtin_bit_cast(v8h, cur[s]), bp = __builtin_bit_cast(v8h, dp);
      asm volatile("" : "+v"(fl_lane));
#pragma unroll
      for (int mt = 0; mt < kSpecMT; ++mt) {
        const v8h fh = __builtin_bit_cast(v8h, F[((mt * kSpecKS + s) * 2 + 0) * 64 + fl_lane]);
        const v8h fl = __builtin_bit_cast(v8h, F[((mt * kSpecKS + s) * 2 + 1) * 64 + fl_lane]);
        are[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fh, b, are[mt], 0, 0, 0);
        aim[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fh, bp, aim[mt], 0, 0, 0);
        are[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fl, b, are[mt], 0, 0, 0);
        aim[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(fl, bp, aim[mt], 0, 0, 0);
      }
    }
#pragma unroll
    for (int mt = 0; mt < kSpecMT; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        sum[mt][i] += __builtin_amdgcn_sqrtf(are[mt][i] * are[mt][i] + aim[mt][i] * aim[mt][i]);
    if (ntl == 0) {
      // row done: reduce-scatter over lane bits 3..0, lane keeps q = lane & 15
      // of the 16 sums a[q] = sum[q >> 2][q & 3]
      float a[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) a[q] = sum[q >> 2][q & 3];
#pragma unroll
      for (int m = 8; m >= 1; m >>= 1) {
        const bool up = (lane & m) != 0;
#pragma unroll
        for (int i = 0; i < m; ++i) {
          const float send = up ? a[i] : a[i + m];
          const float keep = up ? a[i + m] : a[i];
          a[i] = keep + __shfl_xor(send, m, 64);
        }
      }
      const int q = lane & 15;
      const int g = 16 * (q >> 2) + 4 * kq + (q & 3);
      if (g < kSpecOut) out[row * kSpecOut + g] = a[0];
#pragma unroll
      for (int mt = 0; mt < kSpecMT; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) sum[mt][i] = 0.0f;
    }
    row = nrow;
    tile = ntl;
#pragma unroll
    for (int s = 0; s < kSpecKS; ++s) cur[s] = nxt[s];
  }
}

hipError_t launch_spectra(const __half *beams, const void *table, int C, int nbeams, int T, int NT, float *out,
                          hipStream_t s) {
  if (NT != kSpecFFT || T % kSpecFFT || !table) return hipErrorInvalidValue;
  int dev = 0, ncu = 256, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
    ncu = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, spectra_kernel, 64 * kSpecWaves, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int rows = C * 2 * nbeams;
  const int grid = std::min((rows + kSpecWaves - 1) / kSpecWaves, per_cu * ncu);
  hipLaunchKernelGGL(spectra_kernel, dim3(grid), dim3(64 * kSpecWaves), 0, s, reinterpret_cast<const su4 *>(beams),
                     reinterpret_cast<const su4 *>(table), rows, T / kSpecFFT, out);
  return hipGetLastError();
}



// ============================================================ bandpass
// LambdaGPUPipeline's bandpass product (lambda_gpu_pipeline.hpp:522-531,
// 814-816; detect_and_downsample_fft, spatial.cuh:795-835): per (channel,
// pol, beam) the forward FFT over the whole block (N = NP * NT samples), then
// the mean of |X|^2 over each run of D consecutive bins (natural order; NaN
// values are skipped) -> float [C][P][B][N / D].
//
// One workgroup per row, the row in LDS as FP32 complex (N <= 8192: 64 KB),
// a mixed-radix Stockham FFT in place: each pass reads every butterfly's R
// inputs into registers, then (after a barrier) writes the R-point DFTs to
// their autosorted places, so the result is in natural order with no
// bit-reversal pass.  N = 2^r 32^q: one 2^r-point pass, then radix-32
// passes (8192 = 8 x 32 x 32: three passes, bank-conflict free with one pad
// slot per 32 elements; radix 16 + 16 + 16 + 2 took four passes).  Each
// butterfly's twiddles are powers of W_N^{k N/(Ns R)}, computed with sincospif
// (a table read from global memory put its latency into every pass).
namespace {

// cos / sin (2 pi k / 32), k = 0..15: the twiddles inside DFTs of up to 32 points.
__device__ constexpr float kCos32[16] = {
    1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f,
    0.70710678118654752f, 0.55557023301960218f, 0.38268343236508977f, 0.19509032201612826f,
    0.0f, -0.19509032201612826f, -0.38268343236508977f, -0.55557023301960218f,
    -0.70710678118654752f, -0.83146961230254524f, -0.92387953251128674f, -0.98078528040323043f};
__device__ constexpr float kSin32[16] = {
    0.0f, 0.19509032201612826f, 0.38268343236508977f, 0.55557023301960218f,
    0.70710678118654752f, 0.83146961230254524f, 0.92387953251128674f, 0.98078528040323043f,
    1.0f, 0.98078528040323043f, 0.92387953251128674f, 0.83146961230254524f,
    0.70710678118654752f, 0.55557023301960218f, 0.38268343236508977f, 0.19509032201612826f};

// t = W32^I * o for I = 0..15 (W32 = exp(-2 pi i / 32)); 1 and -i need no multiply.
template <int I>
__device__ __forceinline__ float2 w32_mul(float2 o) {
  if constexpr (I == 0) {
    return o;
  } else if constexpr (I == 8) {
    return make_float2(o.y, -o.x);
  } else {
    const float c = kCos32[I], s = kSin32[I];
    return make_float2(c * o.x + s * o.y, c * o.y - s * o.x);
  }
}

template <int R, int K>
__device__ __forceinline__ void dft_combine(float2 *v, const float2 *e, const float2 *o) {
  if constexpr (K < R / 2) {
    const float2 t = w32_mul<K * (32 / R)>(o[K]);
    v[K] = make_float2(e[K].x + t.x, e[K].y + t.y);
    v[K + R / 2] = make_float2(e[K].x - t.x, e[K].y - t.y);
    dft_combine<R, K + 1>(v, e, o);
  }
}

// In-register forward DFT of R = 2^k <= 32 points, natural order in and out
// (radix-2 decimation in time, fully unrolled).
template <int R>
__device__ __forceinline__ void dft_reg(float2 *v) {
  if constexpr (R > 1) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int i = 0; i < R / 2; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    dft_reg<R / 2>(e);
    dft_reg<R / 2>(o);
    dft_combine<R, 0>(v, e, o);
  }
}

constexpr int kBpThreads = 512, kBpMaxN = 8192;

// Sum over the 64 lanes of a wavefront (xor butterflies), all lanes get it.
__device__ __forceinline__ float wave_sum_bp(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// LDS slot of element i: one pad slot per 32, so the strided writes of the
// early passes spread over the banks.
__device__ __forceinline__ int bp_slot(int i) { return i + (i >> 5); }

// One Stockham pass of radix R over N points with completed sub-transform size
// Ns: butterfly j reads x[j + r N/R], twiddles by W_{Ns R}^{(j mod Ns) r}, and
// writes x[(j / Ns) Ns R + (j mod Ns) + r Ns].  The twiddles of a butterfly are
// powers of one table entry (W_N^{k N/(Ns R)}), formed by repeated products.
template <int R>
__device__ __forceinline__ void stockham_pass(float2 *buf, int N, int Ns) {
  constexpr int kG = (kBpMaxN / R + kBpThreads - 1) / kBpThreads;  // butterflies per thread at most
  const int nj = N / R, stride = N / R, tstep = N / (Ns * R);
  float2 v[kG][R];
#pragma unroll
  for (int g = 0; g < kG; ++g) {
    const int j = threadIdx.x + g * kBpThreads;
    if (j < nj) {
      const int k = j % Ns;
      float2 w1;  // W_N^{k tstep}: angle -2 pi k tstep / N, k tstep < N exact in FP32
      sincospif(-2.0f * (float)(k * tstep) / (float)N, &w1.y, &w1.x);
      float2 w = make_float2(1.0f, 0.0f);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float2 x = buf[bp_slot(j + r * stride)];
        v[g][r] = r == 0 ? x : make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
        w = make_float2(w.x * w1.x - w.y * w1.y, w.x * w1.y + w.y * w1.x);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int g = 0; g < kG; ++g) {
    const int j = threadIdx.x + g * kBpThreads;
    if (j < nj) {
      dft_reg<R>(v[g]);
      const int k = j % Ns, base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[bp_slot(base + r * Ns)] = v[g][r];
    }
  }
  __syncthreads();
}

}  // namespace

__global__ void __launch_bounds__(kBpThreads)
bandpass_kernel(const __half2 *__restrict__ beams, int N, int D,
                float *__restrict__ out) {
  extern __shared__ __align__(16) char bp_smem[];
  float2 *buf = reinterpret_cast<float2 *>(bp_smem);
  const int M = N / D;
  float *sums = reinterpret_cast<float *>(buf + bp_slot(N));
  int *cnt = reinterpret_cast<int *>(sums + M);
  const long row = blockIdx.x;
  const __half2 *x = beams + row * N;
  for (int i = threadIdx.x; i < N; i += kBpThreads) buf[bp_slot(i)] = __half22float2(x[i]);
  for (int m = threadIdx.x; m < M; m += kBpThreads) {
    sums[m] = 0.0f;
    cnt[m] = 0;
  }
  __syncthreads();
  // N = 2^m = 2^r 32^q: the 2^r-point pass first (it has the most butterflies
  // and the least register pressure), then q radix-32 passes
  int m = 0;
  while ((1 << m) < N) ++m;
  const int r = m % 5;
  int Ns = 1;
  if (r == 1) stockham_pass<2>(buf, N, Ns);
  else if (r == 2) stockham_pass<4>(buf, N, Ns);
  else if (r == 3) stockham_pass<8>(buf, N, Ns);
  else if (r == 4) stockham_pass<16>(buf, N, Ns);
  Ns <<= r;
  while (Ns < N) {
    stockham_pass<32>(buf, N, Ns);
    Ns *= 32;
  }
  // mean of |X|^2 over runs of D bins, NaN skipped
  const int lane = threadIdx.x & 63;
  for (int i0 = threadIdx.x - lane; i0 < N; i0 += kBpThreads) {
    const int i = i0 + lane;  // a wave covers 64 consecutive bins
    const float2 X = i < N ? buf[bp_slot(i)] : make_float2(0.0f, 0.0f);
    const float p = X.x * X.x + X.y * X.y;
    const bool ok = i < N && !isnan(p);
    if (D % 64 == 0) {  // the wave's 64 bins share one run: one atomic per wave
      const float ws = wave_sum_bp(ok ? p : 0.0f);
      const int wc = __popcll(__ballot(ok));
      if (lane == 0 && i0 < N) {
        atomicAdd(&sums[i0 / D], ws);
        atomicAdd(&cnt[i0 / D], wc);
      }
    } else if (ok) {
      atomicAdd(&sums[i / D], p);
      atomicAdd(&cnt[i / D], 1);
    }
  }
  __syncthreads();
  for (int m = threadIdx.x; m < M; m += kBpThreads) out[row * M + m] = cnt[m] > 0 ? sums[m] / (float)cnt[m] : 0.0f;
}

hipError_t launch_bandpass(const __half *beams, int rows, int N, int D, float *out, hipStream_t s) {
  if (N < 2 || N > kBpMaxN || (N & (N - 1)) || D < 1 || N % D) return hipErrorInvalidValue;
  const size_t sm = (size_t)(N + N / 32) * sizeof(float2) + (size_t)(N / D) * (sizeof(float) + sizeof(int));
  if (sm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(bandpass_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(bandpass_kernel, dim3(rows), dim3(kBpThreads), sm, s, reinterpret_cast<const __half2 *>(beams),
                     N, D, out);
  return hipGetLastError();
}

}  // namespace lam
