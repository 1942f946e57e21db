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
  if (hipGetDevice(&dev) == hipSuccess) hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, spectra_kernel, 64 * kSpecWaves, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int rows = C * 2 * nbeams;
  const int grid = std::min((rows + kSpecWaves - 1) / kSpecWaves, per_cu * ncu);
  hipLaunchKernelGGL(spectra_kernel, dim3(grid), dim3(64 * kSpecWaves), 0, s, reinterpret_cast<const su4 *>(beams),
                     reinterpret_cast<const su4 *>(table), rows, T / kSpecFFT, out);
  return hipGetLastError();
}

}  // namespace lam
