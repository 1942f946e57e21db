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
// One workgroup per (channel, pol, beam); each wave takes every fourth packet,
// lane t holds sample t, and the FFT runs as six radix-2 decimation-in-frequency
// stages across the wavefront (partner value by lane-xor shuffle), leaving
// X[bitrev(t)] in lane t.  Input: the FP16 beams the pipeline writes.
#include "lam_internal.hpp"

namespace lam {

constexpr int kSpecFFT = 64, kSpecDrop = 5, kSpecWaves = 4;

__device__ __forceinline__ int bitrev6(int x) { return (int)(__builtin_bitreverse32((uint32_t)x) >> 26); }

__global__ void __launch_bounds__(64 * kSpecWaves)
spectra_kernel(const __half2 *__restrict__ beams, int nbeams, int T, int np, float *__restrict__ out) {
  __shared__ float part[kSpecWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long row = blockIdx.x;  // (c * 2 + p) * nbeams + e
  const __half2 *x = beams + row * T;
  // twiddle of each stage for this lane: W_{2h}^(lane mod h) = exp(-2 pi i (lane mod h) / (2h))
  float2 tw[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int h = 32 >> s;
    float sn, cs;
    sincospif(-(float)(lane & (h - 1)) / (float)h, &sn, &cs);
    tw[s] = make_float2(cs, sn);
  }
  float acc = 0.0f;
  for (int o = w; o < np; o += kSpecWaves) {
    float2 v = __half22float2(x[(long)o * kSpecFFT + lane]);
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int h = 32 >> s;
      const float2 y = make_float2(__shfl_xor(v.x, h, 64), __shfl_xor(v.y, h, 64));
      if (lane & h) {
        const float2 d = make_float2(y.x - v.x, y.y - v.y);
        v = make_float2(d.x * tw[s].x - d.y * tw[s].y, d.x * tw[s].y + d.y * tw[s].x);
      } else {
        v = make_float2(v.x + y.x, v.y + y.y);
      }
    }
    acc += sqrtf(v.x * v.x + v.y * v.y);
  }
  part[w][lane] = acc;
  __syncthreads();
  if (w == 0) {
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < kSpecWaves; ++q) s += part[q][lane];
    const int u = bitrev6(lane);
    const int keep = kSpecFFT / 2 - kSpecDrop;  // 27 bins each side
    int g = -1;
    if (u >= kSpecFFT - keep) g = u - (kSpecFFT - keep);
    else if (u < keep) g = u + keep;
    if (g >= 0) out[row * (kSpecFFT - 2 * kSpecDrop) + g] = s;
  }
}

hipError_t launch_spectra(const __half *beams, int C, int nbeams, int T, int NT, float *out, hipStream_t s) {
  if (NT != kSpecFFT || T % kSpecFFT) return hipErrorInvalidValue;
  hipLaunchKernelGGL(spectra_kernel, dim3(C * 2 * nbeams), dim3(64 * kSpecWaves), 0, s,
                     reinterpret_cast<const __half2 *>(beams), nbeams, T, T / kSpecFFT, out);
  return hipGetLastError();
}

}  // namespace lam
