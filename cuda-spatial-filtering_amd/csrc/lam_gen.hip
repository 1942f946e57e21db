// Seeded synthetic voltage generator: Gaussian noise from a counter-based
// hash plus CW tones with per-receiver steering phasors, quantised to the
// reference's complex<int8> ring layout (packet_formats.hpp:265-407).
// Benchmark and test input only; nothing in the pipeline depends on it.
#include "lam_internal.hpp"

namespace lam {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Two standard normals from one 64-bit hash (Box-Muller).
__device__ __forceinline__ float2 gauss2(uint64_t h) {
  const float u1 = ((float)((uint32_t)(h >> 40) + 1u)) * (1.0f / 16777216.0f);  // (0, 1]
  const float u2 = ((float)((uint32_t)(h >> 8) & 0xffffffu)) * (1.0f / 16777216.0f);
  const float r = sqrtf(-2.0f * logf(u1));
  float s, c;
  sincospif(2.0f * u2, &s, &c);
  return make_float2(r * c, r * s);
}

__device__ __forceinline__ int8_t quant(float v) {
  return (int8_t)fminf(127.0f, fmaxf(-127.0f, rintf(v)));
}

// One thread per (c, slot, f, t): RPP receivers x 2 pols.
__global__ void __launch_bounds__(256) synth_kernel(GenParams g) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int NS = g.NP + 2;
  const long nthreads = (long)g.C * NS * g.F * g.NT;
  // unit scales [C][NS][R][2]
  const long nscale = (long)g.C * NS * g.R * 2;
  for (long i = gid; i < nscale; i += (long)gridDim.x * blockDim.x) g.scales[i] = 1;
  if (gid >= nthreads) return;
  const int t = (int)(gid % g.NT);
  long rest = gid / g.NT;
  const int f = (int)(rest % g.F);
  rest /= g.F;
  const int slot = (int)(rest % NS);
  const int c = (int)(rest / NS);
  const long tt = (long)(slot - 1) * g.NT + t;  // slot 1 starts the block
  const float gain = g.gain[c];
  // tone phasors of this time step
  float2 ph[kGenMaxSources];
  float2 g1[kGenMaxSources];
  float g0[kGenMaxSources];
  for (int s = 0; s < g.S; ++s) {
    const float *tn = g.tones + ((long)c * g.S + s) * 6;
    const double cyc = (double)tn[1] * (double)tt;
    const float fr = (float)(cyc - floor(cyc));
    float sn, cs;
    sincosf(6.283185307179586f * fr + tn[2], &sn, &cs);
    ph[s] = make_float2(tn[0] * cs, tn[0] * sn);
    g0[s] = tn[3];
    g1[s] = make_float2(tn[4], tn[5]);
  }
  int8_t *out = g.samples + gid * (long)g.RPP * 4;
  const uint64_t base = ((uint64_t)g.seed * 0x9e3779b97f4a7c15ULL) ^
                        ((uint64_t)g.block_index << 44);
  for (int rr = 0; rr < g.RPP; ++rr) {
    const int r = f * g.RPP + rr;
    const uint64_t ctr = (uint64_t)(gid * g.RPP + rr) * 2;
    const float2 n0 = gauss2(mix64(base + ctr)), n1 = gauss2(mix64(base + ctr + 1));
    float2 x0 = make_float2(g.sigma * n0.x, g.sigma * n0.y);
    float2 x1 = make_float2(g.sigma * n1.x, g.sigma * n1.y);
    for (int s = 0; s < g.S; ++s) {
      const float2 a = g.steer[((long)c * g.S + s) * g.R + r];
      const float2 v = make_float2(ph[s].x * a.x - ph[s].y * a.y, ph[s].x * a.y + ph[s].y * a.x);
      x0.x += g0[s] * v.x;
      x0.y += g0[s] * v.y;
      x1.x += g1[s].x * v.x - g1[s].y * v.y;
      x1.y += g1[s].x * v.y + g1[s].y * v.x;
    }
    char4 q;
    q.x = quant(gain * x0.x);
    q.y = quant(gain * x0.y);
    q.z = quant(gain * x1.x);
    q.w = quant(gain * x1.y);
    *reinterpret_cast<char4 *>(out + rr * 4) = q;
  }
}

hipError_t launch_synth(const GenParams &g, hipStream_t s) {
  if (g.S > kGenMaxSources) return hipErrorInvalidValue;
  const long n = (long)g.C * (g.NP + 2) * g.F * g.NT;
  const long nb = (n + 255) / 256;
  hipLaunchKernelGGL(synth_kernel, dim3((unsigned)nb), dim3(256), 0, s, g);
  return hipGetLastError();
}

}  // namespace lam
