// Learned warm start of the eigen stage: feature extraction from the block's
// visibilities, a 2-layer MLP (BF16 weights, FP32 accumulation) on MFMA for
// inference, and the FP32 backward / Adam kernels that train it on the GPU.
//
// Features of one (channel, pol) problem (n receivers, KW warm vectors):
//   [0, n(n-1))            whitened covariance A(i,j)/(d_i d_j), i < j, ordered
//                          idx = j(j-1)/2 + i, (re, im) interleaved
//   [n(n-1), +2 KW n)      previous block's top-KW whitened eigenvectors x sqrt(n)
//   zero padding to Dp (multiple of 32)
// Output: KW complex vectors of n entries (re, im interleaved), padded to Dyp.
// The covariance and whitening follow spatial.cuh:672-713 and 1065-1093.
#include "lam_internal.hpp"

namespace lam {

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---- features: one workgroup per (channel, pol) ----
// Each thread owns PER baselines of the problem's contiguous XX or YY run and
// issues their loads for four splits at a time unconditionally (indices
// clamped, surplus discarded), so the kernel waits on memory about
// nsplit / 4 times; diagonal baselines give the whitening.
constexpr int kFeatThreads = 256;

template <int PER>
__global__ void __launch_bounds__(kFeatThreads) feature_kernel(FeatParams P) {
  __shared__ float d[64];
  const int tid = threadIdx.x, cp = blockIdx.x, c = cp >> 1, p = cp & 1;
  const int n = P.n, nbl = n * (n + 1) / 2;
  const int2 *vis = reinterpret_cast<const int2 *>(P.vis_part) + ((long)c * 4 + 3 * p) * nbl;
  const long split_stride = (long)P.C * 4 * nbl;
  int blc[PER];
  long re[PER], im[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    blc[r] = min(tid + r * kFeatThreads, nbl - 1);
    re[r] = 0;
    im[r] = 0;
  }
  for (int k0 = 0; k0 < P.nsplit; k0 += 4) {
    int2 x[4][PER];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int2 *v = vis + (long)(k0 + u < P.nsplit ? k0 + u : 0) * split_stride;
#pragma unroll
      for (int r = 0; r < PER; ++r) x[u][r] = v[blc[r]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k0 + u >= P.nsplit) break;
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        re[r] += x[u][r].x;
        im[r] += x[u][r].y;
      }
    }
  }
  int ii[PER], jj[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int bl = tid + r * kFeatThreads;
    int j = (int)((sqrtf(8.0f * bl + 1.0f) - 1.0f) * 0.5f);
    while (j * (j + 1) / 2 > bl) --j;
    while ((j + 1) * (j + 2) / 2 <= bl) ++j;
    ii[r] = bl - j * (j + 1) / 2;
    jj[r] = j;
    if (bl < nbl && ii[r] == j) d[j] = sqrtf(fmaxf((float)re[r], 1e-10f));
  }
  __syncthreads();
  uint32_t *x32 = reinterpret_cast<uint32_t *>(P.X + (long)cp * P.Dp);
  // whitened off-diagonal entries: baseline bl = j (j + 1) / 2 + i -> feature e = j (j - 1) / 2 + i
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int bl = tid + r * kFeatThreads, i = ii[r], j = jj[r];
    if (bl < nbl && i != j) {
      const float inv = 1.0f / (d[i] * d[j]);
      x32[j * (j - 1) / 2 + i] = (uint32_t)f2bf((float)re[r] * inv) | ((uint32_t)f2bf((float)im[r] * inv) << 16);
    }
  }
  // previous basis x sqrt(n), then zero padding (Dp and the offsets are even)
  const int boff = n * (n - 1) / 2, nb = P.KW * n;
  const float sq = sqrtf((float)n);
  for (int e = tid; e < nb; e += kFeatThreads) {
    float2 b = make_float2(0.f, 0.f);
    if (P.basis) b = P.basis[(long)cp * nb + e];
    x32[boff + e] = (uint32_t)f2bf(sq * b.x) | ((uint32_t)f2bf(sq * b.y) << 16);
  }
  for (int e = boff + nb + tid; e < P.Dp / 2; e += kFeatThreads) x32[e] = 0u;
}

hipError_t launch_features(const FeatParams &p, hipStream_t s) {
  if (p.Dp % 2 || p.n < 1 || p.n > 64 || p.nsplit < 1) return hipErrorInvalidValue;
  const int per = (p.n * (p.n + 1) / 2 + kFeatThreads - 1) / kFeatThreads;  // 1..9
  const dim3 g(p.C * 2), b(kFeatThreads);
  switch (per) {
#define LAM_FEAT(N) \
  case N: hipLaunchKernelGGL(feature_kernel<N>, g, b, 0, s, p); break;
    LAM_FEAT(1) LAM_FEAT(2) LAM_FEAT(3) LAM_FEAT(4) LAM_FEAT(5) LAM_FEAT(6) LAM_FEAT(7) LAM_FEAT(8) LAM_FEAT(9)
#undef LAM_FEAT
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---- dense layer on MFMA: Z[M][N] = X[M][K] Wt[N][K]^T + b[N] (+ skip) ----
// The output layer adds a skip connection from the previous block's vectors
// held in the features: Z[m][j] += scale * X'[m][off + j] for j < len.
// One workgroup per 32x32 output tile; its NW waves split K in 16-wide steps
// (U steps' operands loaded before their MFMAs, so a wave waits on memory
// about K / (16 NW U) times) and reduce through LDS.  Output: FP32 Z and/or
// BF16 relu(Z).
template <int NW, int U>
__global__ void __launch_bounds__(64 * NW)
dense_fwd_kernel(const uint16_t *__restrict__ X, int M, int K, const uint16_t *__restrict__ Wt,
                 const float *__restrict__ b, int N, float *__restrict__ Z,
                 uint16_t *__restrict__ Hbf, Skip sk) {
  __shared__ float red[NW][16][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
  const int row = lane & 31, kh = lane >> 5;
  const int m = m0 + row;
  const bool mv = m < M;
  const uint16_t *xa = X + (long)(mv ? m : 0) * K + kh * 8;
  const uint16_t *wb = Wt + (long)(n0 + row) * K + kh * 8;
  v16f acc = {};
  const int nks = K / 16;
  int ks = w;
  for (; ks + (U - 1) * NW < nks; ks += U * NW) {
    v8bf a[U], bb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = *reinterpret_cast<const v8bf *>(xa + (ks + NW * u) * 16);
      bb[u] = *reinterpret_cast<const v8bf *>(wb + (ks + NW * u) * 16);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!mv) a[u] = v8bf{};
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[u], bb[u], acc, 0, 0, 0);
    }
  }
  for (; ks < nks; ks += NW) {
    v8bf a = *reinterpret_cast<const v8bf *>(xa + ks * 16);
    const v8bf bb = *reinterpret_cast<const v8bf *>(wb + ks * 16);
    if (!mv) a = v8bf{};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) red[w][v][lane] = acc[v];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16 / NW; ++j) {
    const int o = threadIdx.x + 64 * NW * j;
    const int l = o & 63, v = o >> 6;
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < NW; ++q) s += red[q][v][l];
    const int i = 8 * (v >> 2) + 4 * (l >> 5) + (v & 3), col = l & 31;
    const int mm = m0 + i, nn = n0 + col;
    if (mm < M) {
      float z = s + b[nn];
      if (sk.X && nn < sk.len) z += sk.scale * bf2f(sk.X[(long)mm * sk.ld + sk.off + nn]);
      if (Z) Z[(long)mm * N + nn] = z;
      if (Hbf) Hbf[(long)mm * N + nn] = f2bf(fmaxf(z, 0.0f));
    }
  }
}

hipError_t launch_dense_fwd(const uint16_t *X, int M, int K, const uint16_t *Wt, const float *b,
                            int N, float *Z, uint16_t *Hbf, const Skip &sk, hipStream_t s) {
  if (K % 16 || N % 32) return hipErrorInvalidValue;
  const dim3 grid((M + 31) / 32, N / 32);
  if (K >= 1024)  // long reductions: more waves per tile, deeper loads
    hipLaunchKernelGGL((dense_fwd_kernel<8, 8>), grid, dim3(512), 0, s, X, M, K, Wt, b, N, Z, Hbf, sk);
  else
    hipLaunchKernelGGL((dense_fwd_kernel<4, 4>), grid, dim3(256), 0, s, X, M, K, Wt, b, N, Z, Hbf, sk);
  return hipGetLastError();
}

// ---- inference: the two dense layers ----
// Timed against a layer 1 split over K (FP32 partial tiles) with a layer 2
// that summed them on the fly: 24 us per block at 768 problems against 21 us
// for these two launches, so inference uses the training forward kernels.
hipError_t launch_mlp_infer(const uint16_t *X, int M, int K, const uint16_t *W1, const float *b1, int H,
                            const uint16_t *W2, const float *b2, int N, uint16_t *Hbf, float *Y, const Skip &sk,
                            hipStream_t s) {
  if (hipError_t e = launch_dense_fwd(X, M, K, W1, b1, H, nullptr, Hbf, Skip{}, s)) return e;
  return launch_dense_fwd(Hbf, M, H, W2, b2, N, Y, nullptr, sk, s);
}

// ---- FP32 GEMM for the backward pass ----
// C[i][j] = sum_k A(i,k) B(k,j), A(i,k) = AT ? A[k*lda+i] : A[i*lda+k] (FP32),
// B(k,j) = Bbf[k*ldb+j] (BF16); optional ReLU mask: C *= (mask[i*ldm+j] > 0).
template <bool AT>
__global__ void __launch_bounds__(256)
gemm_bwd_kernel(const float *__restrict__ A, int lda, const uint16_t *__restrict__ Bbf, int ldb,
                float *__restrict__ C, int ldc, int I, int J, int K, const float *__restrict__ mask,
                int ldm) {
  __shared__ float As[16][64 + 4];
  __shared__ float Bs[16][64 + 4];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = tid + 256 * u;  // 0..1023
      int ii, kk;
      if (AT) {
        ii = e & 63;
        kk = e >> 6;
      } else {
        kk = e & 15;
        ii = e >> 4;
      }
      const int gi = i0 + ii, gk = k0 + kk;
      float v = 0.0f;
      if (gi < I && gk < K) v = AT ? A[(long)gk * lda + gi] : A[(long)gi * lda + gk];
      As[kk][ii] = v;
      const int jj = e & 63, kb = e >> 6;
      const int gj = j0 + jj, gkb = k0 + kb;
      Bs[kb][jj] = (gj < J && gkb < K) ? bf2f(Bbf[(long)gkb * ldb + gj]) : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = As[kk][ty * 4 + u];
        bv[u] = Bs[kk][tx * 4 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fmaf(a[u], bv[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int gi = i0 + ty * 4 + u;
    if (gi >= I) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int gj = j0 + tx * 4 + v;
      if (gj >= J) continue;
      float r = acc[u][v];
      if (mask && !(mask[(long)gi * ldm + gj] > 0.0f)) r = 0.0f;
      C[(long)gi * ldc + gj] = r;
    }
  }
}

hipError_t launch_gemm_bwd(bool at, const float *A, int lda, const uint16_t *B, int ldb, float *C,
                           int ldc, int I, int J, int K, const float *mask, int ldm, hipStream_t s) {
  const dim3 grid((J + 63) / 64, (I + 63) / 64);
  if (at)
    hipLaunchKernelGGL(gemm_bwd_kernel<true>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, I, J,
                       K, mask, ldm);
  else
    hipLaunchKernelGGL(gemm_bwd_kernel<false>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, I, J,
                       K, mask, ldm);
  return hipGetLastError();
}

// db[j] = sum_m D[m][j]
__global__ void colsum_kernel(const float *__restrict__ D, int M, int N, float *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.0f;
  for (int m = 0; m < M; ++m) s += D[(long)m * N + j];
  out[j] = s;
}

hipError_t launch_colsum(const float *D, int M, int N, float *out, hipStream_t s) {
  hipLaunchKernelGGL(colsum_kernel, dim3((N + 255) / 256), dim3(256), 0, s, D, M, N, out);
  return hipGetLastError();
}

// ---- phase-invariant subspace loss, one wave per sample ----
// L = (1/M) sum_m sum_{k < K_m} (1 - |u_k^H y_k|^2 / (|y_k|^2 + eps))
// dL/dRe(y) + i dL/dIm(y) = -(2/M) (s u - f y) / (|y|^2 + eps),  s = u^H y, f = |s|^2 / (|y|^2 + eps)
__global__ void __launch_bounds__(256)
loss_kernel(const float *__restrict__ Y, int Dyp, const float2 *__restrict__ U,
            const int *__restrict__ Kl, int M, int n, int KW, float *__restrict__ dY,
            float *__restrict__ loss_acc) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= M) return;
  const float eps = 1e-12f;
  const float invM = 1.0f / (float)M;
  const int K = min(Kl[m], KW);
  const float2 *y = reinterpret_cast<const float2 *>(Y + (long)m * Dyp);
  float2 *dy = reinterpret_cast<float2 *>(dY + (long)m * Dyp);
  float total = 0.0f;
  for (int k = 0; k < KW; ++k) {
    float sr = 0.f, si = 0.f, nn = 0.f;
    for (int i = lane; i < n; i += 64) {
      const float2 yv = y[k * n + i], uv = U[((long)m * KW + k) * n + i];
      sr += uv.x * yv.x + uv.y * yv.y;  // conj(u) y
      si += uv.x * yv.y - uv.y * yv.x;
      nn += yv.x * yv.x + yv.y * yv.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      sr += __shfl_xor(sr, o, 64);
      si += __shfl_xor(si, o, 64);
      nn += __shfl_xor(nn, o, 64);
    }
    const float den = nn + eps;
    const float f = (sr * sr + si * si) / den;
    const bool on = k < K;
    if (on) total += 1.0f - f;
    for (int i = lane; i < n; i += 64) {
      float2 g = make_float2(0.f, 0.f);
      if (on) {
        const float2 yv = y[k * n + i], uv = U[((long)m * KW + k) * n + i];
        // s u - f y with s = sr + i si
        const float gr = (sr * uv.x - si * uv.y) - f * yv.x;
        const float gi = (sr * uv.y + si * uv.x) - f * yv.y;
        const float sc = -2.0f * invM / den;
        g = make_float2(sc * gr, sc * gi);
      }
      dy[k * n + i] = g;
    }
  }
  // padded outputs carry no gradient
  for (int e = 2 * KW * n + lane; e < Dyp; e += 64) dY[(long)m * Dyp + e] = 0.0f;
  if (lane == 0 && loss_acc) atomicAdd(loss_acc, total * invM);
}

hipError_t launch_loss(const float *Y, int Dyp, const float2 *U, const int *K, int M, int n,
                       int KW, float *dY, float *loss_acc, hipStream_t s) {
  hipLaunchKernelGGL(loss_kernel, dim3((M + 3) / 4), dim3(256), 0, s, Y, Dyp, U, K, M, n, KW, dY,
                     loss_acc);
  return hipGetLastError();
}

// ---- Adam on the FP32 master parameters, refreshing the BF16 weight copies ----
__global__ void adam_kernel(float *__restrict__ prm, const float *__restrict__ g,
                            float *__restrict__ m1, float *__restrict__ m2, long P, float lr,
                            float b1, float b2, float eps, float c1, float c2, long w1_off,
                            long w1_len, uint16_t *__restrict__ w1bf, long w2_off, long w2_len,
                            uint16_t *__restrict__ w2bf) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  const float mh = a / c1, vh = v / c2;
  const float w = prm[i] - lr * mh / (sqrtf(vh) + eps);
  prm[i] = w;
  if (i >= w1_off && i < w1_off + w1_len) w1bf[i - w1_off] = f2bf(w);
  if (i >= w2_off && i < w2_off + w2_len) w2bf[i - w2_off] = f2bf(w);
}

hipError_t launch_adam(float *prm, const float *g, float *m1, float *m2, long P, float lr, float b1,
                       float b2, float eps, int t, long w1_off, long w1_len, uint16_t *w1bf,
                       long w2_off, long w2_len, uint16_t *w2bf, hipStream_t s) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, prm, g, m1,
                     m2, P, lr, b1, b2, eps, c1, c2, w1_off, w1_len, w1bf, w2_off, w2_len, w2bf);
  return hipGetLastError();
}

// BF16 copies of the weight blocks of the master parameters.
__global__ void to_bf16_kernel(const float *__restrict__ src, long n, uint16_t *__restrict__ dst) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f2bf(src[i]);
}

hipError_t launch_to_bf16(const float *src, long n, uint16_t *dst, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, n,
                     dst);
  return hipGetLastError();
}

// Gather training rows: X[b] = DX[idx[b]], U[b] = DU[idx[b]], K[b] = DK[idx[b]].
__global__ void gather_kernel(const int *__restrict__ idx, int nb, const uint16_t *__restrict__ DX,
                              int Dp, const float2 *__restrict__ DU, int nu,
                              const int *__restrict__ DK, uint16_t *__restrict__ X,
                              float2 *__restrict__ U, int *__restrict__ K) {
  const int b = blockIdx.x;
  if (b >= nb) return;
  const long src = idx[b];
  for (int e = threadIdx.x; e < Dp; e += blockDim.x) X[(long)b * Dp + e] = DX[src * Dp + e];
  for (int e = threadIdx.x; e < nu; e += blockDim.x) U[(long)b * nu + e] = DU[src * nu + e];
  if (threadIdx.x == 0) K[b] = DK[src];
}

hipError_t launch_gather(const int *idx, int nb, const uint16_t *DX, int Dp, const float2 *DU,
                         int nu, const int *DK, uint16_t *X, float2 *U, int *K, hipStream_t s) {
  hipLaunchKernelGGL(gather_kernel, dim3(nb), dim3(256), 0, s, idx, nb, DX, Dp, DU, nu, DK, X, U, K);
  return hipGetLastError();
}

}  // namespace lam
