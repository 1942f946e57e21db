// Small kernels around the hot path: beam-weight operand preparation, the
// visibility dump (accumulator -> float landing layout, then reset) and a
// graph-safe fill kernel.
#include "lam_internal.hpp"

namespace lam {

// Weight operands of B beams for the beam MFMA (corr-beam mode): replaces the
// reference's per-block weightsInputToCCGLIB permutation
// (lambda_corr_beam_only_gpu_pipeline.hpp:332-334).  Effective weight
// w_eff = w * gain * exp(i phi) (phase of the fine delay at channel centre),
// split into hi + lo FP16.
__global__ void prep_weights_kernel(const __half *w, const float2 *gains, const float2 *phase,
                                    int C, int B, int R, int NMT, int NK, int wprec, v8h *wop) {
  const long total = (long)C * 2 * NMT * NK * 64;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int lane = idx & 63;
    long r = idx >> 6;
    const int j = r % NK;
    r /= NK;
    const int mt = r % NMT;
    r /= NMT;
    const int p = r & 1;
    const int c = (int)(r >> 1);
    const int m = 32 * mt + (lane & 31), kg = lane >> 5;
    const int e = m >> 1, u = m & 1;
    v8h hi, lo;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int a = 8 * j + 4 * kg + i;
      float2 ww = make_float2(0.f, 0.f);
      if (e < B && a < R) {
        const __half2 h = reinterpret_cast<const __half2 *>(w)[(((long)c * 2 + p) * B + e) * R + a];
        ww = __half22float2(h);
        const float2 g = gains[((long)c * 2 + p) * R + a];
        const float2 ph = phase[(long)c * R + a];
        float2 t = make_float2(ww.x * g.x - ww.y * g.y, ww.x * g.y + ww.y * g.x);
        ww = make_float2(t.x * ph.x - t.y * ph.y, t.x * ph.y + t.y * ph.x);
      }
      const float v0 = u == 0 ? ww.x : ww.y;
      const float v1 = u == 0 ? -ww.y : ww.x;
      _Float16 h0, l0, h1, l1;
      weight_operands(v0, wprec, h0, l0);
      weight_operands(v1, wprec, h1, l1);
      hi[2 * i] = h0;
      lo[2 * i] = l0;
      hi[2 * i + 1] = h1;
      lo[2 * i + 1] = l1;
    }
    v8h *dst = wop + ((((long)c * 2 + p) * NMT + mt) * NK + j) * 2 * 64;
    dst[lane] = hi;
    dst[64 + lane] = lo;
  }
}

hipError_t launch_prep_weights(const __half *w, const float2 *gains, const float2 *phase,
                               int C, int B, int R, int NMT, int NK, int wprec, v8h *wop,
                               hipStream_t s) {
  const long total = (long)C * 2 * NMT * NK * 64;
  const int blocks = (int)((total + 255) / 256);
  hipLaunchKernelGGL(prep_weights_kernel, dim3(blocks), dim3(256), 0, s, w, gains, phase, C, B,
                     R, NMT, NK, wprec, wop);
  return hipGetLastError();
}

// Accumulator dump (dump_visibilities, lambda_corr_beam_only_gpu_pipeline.hpp:660-701):
// float out[c][bl(i,j)][p][q] = g_{p,i} conj(g_{q,j}) e^{i(phi_i - phi_j)} V_raw,
// optional raw int64 copy, then the accumulator is cleared.
__global__ void dump_vis_kernel(long long *acc, int C, int R, const float2 *gains,
                                const float2 *phase, float *out_f32, long long *out_i64) {
  const long nbl = (long)R * (R + 1) / 2;
  const long total = (long)C * nbl * 4;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int q = idx & 1, p = (idx >> 1) & 1;
    const long cb = idx >> 2;
    const long bl = cb % nbl;
    const int c = (int)(cb / nbl);
    int j = (int)((sqrtf(8.0f * (float)bl + 1.0f) - 1.0f) * 0.5f);
    while ((long)j * (j + 1) / 2 > bl) --j;
    while ((long)(j + 1) * (j + 2) / 2 <= bl) ++j;
    const int i = (int)(bl - (long)j * (j + 1) / 2);
    const long long vr = acc[2 * idx], vi = acc[2 * idx + 1];
    if (out_i64) {
      out_i64[2 * idx] = vr;
      out_i64[2 * idx + 1] = vi;
    }
    if (out_f32) {
      const float2 gi = gains[((long)c * 2 + p) * R + i];
      const float2 gj = gains[((long)c * 2 + q) * R + j];
      const float2 pi = phase[(long)c * R + i], pj = phase[(long)c * R + j];
      // f = gi * conj(gj) * pi * conj(pj)
      float2 f = make_float2(gi.x * gj.x + gi.y * gj.y, gi.y * gj.x - gi.x * gj.y);
      const float2 ph = make_float2(pi.x * pj.x + pi.y * pj.y, pi.y * pj.x - pi.x * pj.y);
      f = make_float2(f.x * ph.x - f.y * ph.y, f.x * ph.y + f.y * ph.x);
      const double xr = (double)vr, xi = (double)vi;
      out_f32[2 * idx] = (float)(xr * f.x - xi * f.y);
      out_f32[2 * idx + 1] = (float)(xr * f.y + xi * f.x);
    }
    acc[2 * idx] = 0;
    acc[2 * idx + 1] = 0;
  }
}

hipError_t launch_dump_vis(long long *acc, int C, int R, const float2 *gains,
                           const float2 *phase, float *out_f32, long long *out_i64,
                           hipStream_t s) {
  const long total = (long)C * R * (R + 1) / 2 * 4;
  const int blocks = (int)std::min<long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(dump_vis_kernel, dim3(blocks), dim3(256), 0, s, acc, C, R, gains, phase,
                     out_f32, out_i64);
  return hipGetLastError();
}

// Graph-safe clear (a kernel rather than hipMemsetAsync).
__global__ void zero_kernel(uint4 *p, size_t n16, unsigned char *tail, size_t ntail) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
  if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  const size_t n16 = bytes / 16;
  unsigned char *tail = reinterpret_cast<unsigned char *>(p) + n16 * 16;
  const int blocks = (int)std::min<size_t>((n16 + 255) / 256 + 1, 8192);
  hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, s, reinterpret_cast<uint4 *>(p),
                     n16, tail, bytes - n16 * 16);
  return hipGetLastError();
}

}  // namespace lam
