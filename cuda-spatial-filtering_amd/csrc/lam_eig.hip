// Eigen / RFI-null stage: one workgroup per (channel, pol) covariance, the
// whole problem resident in LDS.
//
// Replaces, per block, the reference sequence (adaptive pipeline 648-806):
//   visCorrToBaseline + trims + unpack_triangular_baseline_batch (spatial.cuh:672-713)
//   -> diagonalWhiten (spatial.cuh:1065-1093) -> cusolverDnXsyevBatched (adaptive 679-686)
//   -> detectSignalEigenmodeCounts (spatial.cuh:900-950)
//   -> unwhitenEigenvectors + buildIdentityMinusProjection (spatial.cuh:993-1034, 1095-1148)
//      or computeShrinkScaleFactors + scaleEigenvectorColumns + cublasGemmEx loop
//      (spatial.cuh:952-991, 1036-1063; adaptive 726-756)
//   -> gemm_weight_projection W' = W P and the beam-major weight permutes (adaptive 773-798)
// with a single launch.  The eigensolver is a parallel-order (round-robin)
// cyclic Jacobi method: each round applies n/2 disjoint complex rotations,
// rotation angles are computed in FP64 from the FP32 matrix.
#include "lam_internal.hpp"

namespace lam {

struct Rot {
  float c, s, er, ei;  // J = [[c, s], [-s e^{-i phi}, c e^{-i phi}]], e^{i phi} = er + i ei
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// Round-robin position of slot j in round r (circle method, ne even).
__device__ __forceinline__ int rr_pos(int j, int r, int ne) {
  return j == 0 ? 0 : 1 + (j - 1 + r) % (ne - 1);
}

__device__ float block_sum(float v, float *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float t = 0.0f;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

// 2x2 block of a Hermitian update: Y = J_P^H X J_Q with X = A[{p1,q1}][{p2,q2}].
__device__ __forceinline__ void rotate_block(const Rot &RP, const Rot &RQ, float2 &x00, float2 &x01,
                                             float2 &x10, float2 &x11) {
  {  // columns: [x_p, x_q] J_Q
    const float2 eb = make_float2(RQ.er, -RQ.ei);
    const float2 y0 = cmul(eb, x01), y1 = cmul(eb, x11);
    const float2 n00 = make_float2(RQ.c * x00.x - RQ.s * y0.x, RQ.c * x00.y - RQ.s * y0.y);
    const float2 n01 = make_float2(RQ.s * x00.x + RQ.c * y0.x, RQ.s * x00.y + RQ.c * y0.y);
    const float2 n10 = make_float2(RQ.c * x10.x - RQ.s * y1.x, RQ.c * x10.y - RQ.s * y1.y);
    const float2 n11 = make_float2(RQ.s * x10.x + RQ.c * y1.x, RQ.s * x10.y + RQ.c * y1.y);
    x00 = n00; x01 = n01; x10 = n10; x11 = n11;
  }
  {  // rows: J_P^H [x_p; x_q]
    const float2 e = make_float2(RP.er, RP.ei);
    const float2 y0 = cmul(e, x10), y1 = cmul(e, x11);
    const float2 n00 = make_float2(RP.c * x00.x - RP.s * y0.x, RP.c * x00.y - RP.s * y0.y);
    const float2 n01 = make_float2(RP.c * x01.x - RP.s * y1.x, RP.c * x01.y - RP.s * y1.y);
    const float2 n10 = make_float2(RP.s * x00.x + RP.c * y0.x, RP.s * x00.y + RP.c * y0.y);
    const float2 n11 = make_float2(RP.s * x01.x + RP.c * y1.x, RP.s * x01.y + RP.c * y1.y);
    x00 = n00; x01 = n01; x10 = n10; x11 = n11;
  }
}

// FP64 reciprocal square root / reciprocal: hardware estimate plus one Newton step.
__device__ __forceinline__ double rsq64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  return y * (1.5 - 0.5 * x * y * y);
}
__device__ __forceinline__ double rcp64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  return y * (2.0 - x * y);
}

// Rotation that annihilates a_pq (angles in FP64):
// J = [[c, s], [-s e^{-i phi}, c e^{-i phi}]], a_pq = |b| e^{i phi},
// tau = (a_qq - a_pp) / (2|b|), t = sgn(tau) / (|tau| + sqrt(1 + tau^2)).
__device__ __forceinline__ Rot make_rotation(float2 b, float app, float aqq) {
  Rot R{1.0f, 0.0f, 1.0f, 0.0f};
  const double br = b.x, bi = b.y;
  const double b2 = br * br + bi * bi;
  if (b2 > 0.0 && b2 > 1e-60 * ((double)app * app + (double)aqq * aqq)) {
    const double inv_b = rsq64(b2);
    const double tau = ((double)aqq - (double)app) * 0.5 * inv_b;
    const double at = fabs(tau);
    const double u = 1.0 + tau * tau;
    const double t = rcp64(at + u * rsq64(u));
    const double cc = rsq64(1.0 + t * t);
    R.c = (float)cc;
    R.s = (float)((tau >= 0.0 ? t : -t) * cc);
    R.er = (float)(br * inv_b);
    R.ei = (float)(bi * inv_b);
  }
  return R;
}

// Cyclic Jacobi on the Hermitian matrix A (column-major, leading dim n) in LDS;
// V receives the eigenvectors (columns).  Each round rotates n/2 disjoint pairs
// (round-robin order); the update visits only blocks kp <= kq and writes the
// mirrored block as its conjugate transpose.  Returns the number of sweeps.
__device__ int jacobi_lds(float2 *A, float2 *V, int n, int max_sweeps, float tol,
                          int2 *rpq, Rot *rot, float *red, bool init_v = true) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int ne = n + (n & 1), h = ne / 2;
  const int nblk = h * (h + 1) / 2;
  const float inv_n = 1.0f / (float)n;
  for (int i = tid; i < n * n; i += nthr) {
    const int c = (int)(((float)i + 0.5f) * inv_n), r = i - c * n;
    if (init_v) V[i] = make_float2(r == c ? 1.0f : 0.0f, 0.0f);
    if (r == c) A[i].y = 0.0f;
  }
  __syncthreads();
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    float off = 0.0f, tot = 0.0f;
    for (int i = tid; i < n * n; i += nthr) {
      const int c = (int)(((float)i + 0.5f) * inv_n), r = i - c * n;
      const float2 a = A[i];
      const float m = a.x * a.x + a.y * a.y;
      tot += m;
      if (r != c) off += m;
    }
    off = block_sum(off, red);
    tot = block_sum(tot, red);
    if (off == 0.0f || off <= tol * tol * tot) break;
    for (int r = 0; r < ne - 1; ++r) {
      if (tid < h) {
        int p = rr_pos(tid, r, ne), q = rr_pos(ne - 1 - tid, r, ne);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        Rot R{1.0f, 0.0f, 1.0f, 0.0f};
        if (q < n) R = make_rotation(A[q * n + p], A[p * n + p].x, A[q * n + q].x);
        rpq[tid] = make_int2(p, q);
        rot[tid] = R;
      }
      __syncthreads();
      // A <- J^H A J on blocks (kp <= kq); the (kq, kp) block is the conjugate transpose
      for (int idx = tid; idx < nblk; idx += nthr) {
        // idx -> (kp, kq) over the upper triangle, row-major: kq >= kp
        int kp = (int)((2.0f * h + 1.0f - sqrtf((2.0f * h + 1.0f) * (2.0f * h + 1.0f) - 8.0f * idx)) * 0.5f);
        while (kp > 0 && kp * (2 * h - kp + 1) / 2 > idx) --kp;
        while ((kp + 1) * (2 * h - kp) / 2 <= idx) ++kp;
        const int kq = kp + idx - kp * (2 * h - kp + 1) / 2;
        const int2 P = rpq[kp], Q = rpq[kq];
        const Rot RP = rot[kp], RQ = rot[kq];
        const bool pv = P.y < n, qv = Q.y < n;  // the second index of a pair may be the dummy
        float2 x00 = A[Q.x * n + P.x];
        float2 x01 = qv ? A[Q.y * n + P.x] : make_float2(0.f, 0.f);
        float2 x10 = pv ? A[Q.x * n + P.y] : make_float2(0.f, 0.f);
        float2 x11 = (pv && qv) ? A[Q.y * n + P.y] : make_float2(0.f, 0.f);
        rotate_block(RP, RQ, x00, x01, x10, x11);
        if (kp == kq) {  // the annihilated pair: exact zeros, real diagonal
          x01 = make_float2(0.f, 0.f);
          x10 = make_float2(0.f, 0.f);
          x00.y = 0.f;
          x11.y = 0.f;
          A[P.x * n + P.x] = x00;
          if (pv) {
            A[P.y * n + P.x] = x01;
            A[P.x * n + P.y] = x10;
            A[P.y * n + P.y] = x11;
          }
        } else {
          A[Q.x * n + P.x] = x00;
          A[P.x * n + Q.x] = make_float2(x00.x, -x00.y);
          if (qv) {
            A[Q.y * n + P.x] = x01;
            A[P.x * n + Q.y] = make_float2(x01.x, -x01.y);
          }
          if (pv) {
            A[Q.x * n + P.y] = x10;
            A[P.y * n + Q.x] = make_float2(x10.x, -x10.y);
          }
          if (pv && qv) {
            A[Q.y * n + P.y] = x11;
            A[P.y * n + Q.y] = make_float2(x11.x, -x11.y);
          }
        }
      }
      // V <- V J
      const float inv_nn = inv_n;
      for (int idx = tid; idx < n * h; idx += nthr) {
        const int k = (int)(((float)idx + 0.5f) * inv_nn), row = idx - k * n;
        const int2 Q = rpq[k];
        if (Q.y >= n) continue;
        const Rot RQ = rot[k];
        const float2 vp = V[Q.x * n + row], vq = V[Q.y * n + row];
        const float2 y = cmul(make_float2(RQ.er, -RQ.ei), vq);
        V[Q.x * n + row] = make_float2(RQ.c * vp.x - RQ.s * y.x, RQ.c * vp.y - RQ.s * y.y);
        V[Q.y * n + row] = make_float2(RQ.s * vp.x + RQ.c * y.x, RQ.s * vp.y + RQ.c * y.y);
      }
      __syncthreads();
    }
  }
  return sweep;
}

// Ascending order of the diagonal of A: perm[rank] = column, ev[rank] = value.
__device__ void sort_eigen(const float2 *A, int n, int *perm, float *ev) {
  const int j = threadIdx.x;
  if (j < n) {
    const float lj = A[j * n + j].x;
    int rank = 0;
    for (int k = 0; k < n; ++k) {
      const float lk = A[k * n + k].x;
      rank += (lk < lj) || (lk == lj && k < j);
    }
    perm[rank] = j;
    ev[rank] = lj;
  }
  __syncthreads();
}

// Linear-interpolated percentile (pipeline/common.hpp:461-478), FP32 like the reference.
__device__ float pct_linear(const float *v, int n, float q) {
  if (n <= 0) return 0.0f;
  if (n == 1) return v[0];
  const float pos = fminf(1.0f, fmaxf(0.0f, q)) * (float)(n - 1);
  const int lo = (int)floorf(pos), hi = (int)ceilf(pos);
  if (lo == hi) return v[lo];
  return v[lo] + (pos - (float)lo) * (v[hi] - v[lo]);
}

struct EigSmem {
  float2 *A, *V, *W, *G;
  float *ev, *d, *f, *red;
  int *perm;
  int2 *rpq;
  Rot *rot;
  int *kdet;
};

__device__ EigSmem carve(char *base, int n, int B) {
  EigSmem s;
  s.A = reinterpret_cast<float2 *>(base);
  s.V = s.A + n * n;
  s.W = s.V + n * n;
  s.G = s.W + 2 * B * n;
  float *fp = reinterpret_cast<float *>(s.G + B * n);
  s.ev = fp;
  s.d = fp + 64;
  s.f = fp + 128;
  s.red = fp + 192;
  s.perm = reinterpret_cast<int *>(fp + 208);
  s.rpq = reinterpret_cast<int2 *>(s.perm + 64);
  s.rot = reinterpret_cast<Rot *>(s.rpq + 32);
  s.kdet = reinterpret_cast<int *>(s.rot + 32);
  return s;
}

size_t eig_smem_bytes(int n, int B) {
  return sizeof(float2) * (2 * n * n + 3 * B * n) + sizeof(float) * 208 + sizeof(int) * 64 +
         sizeof(int2) * 32 + sizeof(Rot) * 32 + 16;
}

__global__ void __launch_bounds__(kEigThreads) eig_null_kernel(EigParams P) {
  extern __shared__ __align__(16) char smem_raw[];
  const int n = P.R, B = P.B;
  EigSmem S = carve(smem_raw, n, B);
  const int c = blockIdx.x >> 1, p = blockIdx.x & 1;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long nbl = (long)n * (n + 1) / 2;

  // 1. dense Hermitian covariance of pol p (A(i,j) = V[bl(i,j)][p][p], i <= j)
  for (int idx = tid; idx < nbl; idx += nthr) {
    int j = (int)((sqrtf(8.0f * idx + 1.0f) - 1.0f) * 0.5f);
    while ((long)j * (j + 1) / 2 > idx) --j;
    while ((long)(j + 1) * (j + 2) / 2 <= idx) ++j;
    const int i = idx - j * (j + 1) / 2;
    long re = 0, im = 0;
    for (int k = 0; k < P.nsplit; ++k) {
      const int *v = P.vis_part + ((((long)k * P.C + c) * nbl + idx) * 4 + p * 2 + p) * 2;
      re += v[0];
      im += v[1];
    }
    const float2 a = make_float2((float)re, (float)im);
    S.A[j * n + i] = a;
    S.A[i * n + j] = make_float2(a.x, -a.y);
  }
  __syncthreads();
  // 2. whitening (spatial.cuh:1065-1093)
  if (tid < n) S.d[tid] = sqrtf(fmaxf(S.A[tid * n + tid].x, 1e-10f));
  __syncthreads();
  for (int i = tid; i < n * n; i += nthr) {
    const float den = S.d[i % n] * S.d[i / n];
    S.A[i].x /= den;
    S.A[i].y /= den;
  }
  __syncthreads();
  // 3. eigen decomposition
  const int sweeps = jacobi_lds(S.A, S.V, n, P.max_sweeps, P.tol, S.rpq, S.rot, S.red);
  sort_eigen(S.A, n, S.perm, S.ev);
  // 4. K
  if (tid == 0) {
    int k = P.fixed_k[c * 2 + p];
    if (P.detect) {
      const float p20 = pct_linear(S.ev, n, 0.2f), p50 = pct_linear(S.ev, n, 0.5f),
                  p80 = pct_linear(S.ev, n, 0.8f);
      const float thr = p50 + P.delta * ((p80 - p20) / (2.0f * 0.8416f));
      k = 0;
      for (int i = 0; i < n; ++i) k += S.ev[i] > thr;
    }
    S.kdet[0] = max(0, min(n, k));
  }
  __syncthreads();
  const int K = S.kdet[0];
  // 5. output basis U (sorted columns) into S.A: unwhitened+normalised (null) or whitened (shrink)
  if (tid < n) {
    const int src = S.perm[tid];
    float nrm = 0.0f;
    for (int i = 0; i < n; ++i) {
      float2 v = S.V[src * n + i];
      if (!P.shrink) {
        v.x *= S.d[i];
        v.y *= S.d[i];
      }
      S.A[tid * n + i] = v;
      nrm += v.x * v.x + v.y * v.y;
    }
    if (!P.shrink) {
      const float inv = nrm > 0.0f ? rsqrtf(nrm) : 0.0f;
      for (int i = 0; i < n; ++i) {
        S.A[tid * n + i].x *= inv;
        S.A[tid * n + i].y *= inv;
      }
    }
    // per-mode factor: null -> -1 on the K largest, shrink -> lambda_bar / lambda_k
    float fk;
    if (!P.shrink) {
      fk = (tid >= n - K) ? -1.0f : 0.0f;
    } else {
      float lbar = 0.0f;
      if (K < n) {
        for (int i = 0; i < n - K; ++i) lbar += S.ev[i];
        lbar /= (float)(n - K);
      }
      fk = (tid < n - K) ? 1.0f : (S.ev[tid] > 0.0f ? lbar / S.ev[tid] : 0.0f);
    }
    S.f[tid] = fk;
  }
  // weights of this (c, p): rows 0..B-1 original
  for (int idx = tid; idx < B * n; idx += nthr) {
    const __half2 w = reinterpret_cast<const __half2 *>(P.weights)[((long)(c * 2 + p) * B) * n + idx];
    S.W[idx] = __half22float2(w);
  }
  __syncthreads();
  // 6. eigen outputs
  const long cp = (long)c * 2 + p;
  if (tid < n) P.evals[cp * n + tid] = S.ev[tid];
  for (int i = tid; i < n * n; i += nthr)
    reinterpret_cast<float2 *>(P.evecs)[cp * n * n + i] = S.A[i];
  if (tid == 0) {
    P.counts[cp] = K;
    if (P.sweeps) P.sweeps[cp] = sweeps;
  }
  // 7. W' = W P with P = I - U_K U_K^H (null) or V diag(f) V^H (shrink):
  //    G(b,k) = sum_i W(b,i) u_k(i);  W'(b,i) = [null] W(b,i) + sum_k f_k G(b,k) conj(u_k(i))
  for (int idx = tid; idx < B * n; idx += nthr) {
    const int b = idx / n, k = idx % n;
    float2 g = make_float2(0.f, 0.f);
    if (S.f[k] != 0.0f) {
      for (int i = 0; i < n; ++i) {
        const float2 pr = cmul(S.W[b * n + i], S.A[k * n + i]);
        g.x += pr.x;
        g.y += pr.y;
      }
    }
    S.G[b * n + k] = make_float2(g.x * S.f[k], g.y * S.f[k]);
  }
  __syncthreads();
  for (int idx = tid; idx < B * n; idx += nthr) {
    const int b = idx / n, i = idx % n;
    float2 acc = P.shrink ? make_float2(0.f, 0.f) : S.W[b * n + i];
    for (int k = 0; k < n; ++k) {
      const float2 t = cmulc(S.G[b * n + k], S.A[k * n + i]);
      acc.x += t.x;
      acc.y += t.y;
    }
    S.W[(B + b) * n + i] = acc;
  }
  __syncthreads();
  // 8. MFMA operands of the 2B beam rows [W; W'] (split hi + lo FP16)
  const int nrow = 2 * B;
  for (int idx = tid; idx < P.NMT * P.NK * 64; idx += nthr) {
    const int lane = idx & 63, j = (idx >> 6) % P.NK, mt = (idx >> 6) / P.NK;
    const int m = 32 * mt + (lane & 31), kg = lane >> 5;
    const int e = m >> 1, u = m & 1;
    v8h hi, lo;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int a = 8 * j + 4 * kg + i;
      float2 w = make_float2(0.f, 0.f);
      if (e < nrow && a < n) w = S.W[e * n + a];
      const float v0 = u == 0 ? w.x : w.y;
      const float v1 = u == 0 ? -w.y : w.x;
      const _Float16 h0 = (_Float16)v0, h1 = (_Float16)v1;
      hi[2 * i] = h0;
      hi[2 * i + 1] = h1;
      lo[2 * i] = (_Float16)(v0 - (float)h0);
      lo[2 * i + 1] = (_Float16)(v1 - (float)h1);
    }
    v8h *dst = P.wop + ((cp * P.NMT + mt) * P.NK + j) * 2 * 64;
    dst[lane] = hi;
    dst[64 + lane] = lo;
  }
}

hipError_t launch_eig(const EigParams &p, hipStream_t s) {
  const size_t sm = eig_smem_bytes(p.R, p.B);
  if (sm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(eig_null_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(eig_null_kernel, dim3(p.C * 2), dim3(kEigThreads), sm, s, p);
  return hipGetLastError();
}

// ---- standalone batched eigensolver (unit-test entry point) ----
__global__ void __launch_bounds__(kEigThreads)
eigh_batched_kernel(const float2 *Ain, int n, int whiten, int max_sweeps, float tol,
                    float *evals, float2 *evecs, int *sweeps_out) {
  extern __shared__ __align__(16) char smem_raw[];
  EigSmem S = carve(smem_raw, n, 0);
  const int m = blockIdx.x, tid = threadIdx.x;
  for (int i = tid; i < n * n; i += blockDim.x) S.A[i] = Ain[(long)m * n * n + i];
  __syncthreads();
  if (whiten) {
    if (tid < n) S.d[tid] = sqrtf(fmaxf(S.A[tid * n + tid].x, 1e-10f));
    __syncthreads();
    for (int i = tid; i < n * n; i += blockDim.x) {
      const float den = S.d[i % n] * S.d[i / n];
      S.A[i].x /= den;
      S.A[i].y /= den;
    }
    __syncthreads();
  }
  const int sw = jacobi_lds(S.A, S.V, n, max_sweeps, tol, S.rpq, S.rot, S.red);
  sort_eigen(S.A, n, S.perm, S.ev);
  if (tid < n) evals[(long)m * n + tid] = S.ev[tid];
  for (int i = tid; i < n * n; i += blockDim.x) {
    const int k = i / n, r = i % n;
    evecs[(long)m * n * n + i] = S.V[S.perm[k] * n + r];
  }
  if (tid == 0 && sweeps_out) sweeps_out[m] = sw;
}

hipError_t launch_eigh_batched(const float *A, int n, int nmat, int whiten, int max_sweeps,
                               float tol, float *evals, float *evecs, int *sweeps,
                               hipStream_t s) {
  const size_t sm = eig_smem_bytes(n, 0);
  if (sm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(eigh_batched_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(eigh_batched_kernel, dim3(nmat), dim3(kEigThreads), sm, s,
                     reinterpret_cast<const float2 *>(A), n, whiten, max_sweeps, tol, evals,
                     reinterpret_cast<float2 *>(evecs), sweeps);
  return hipGetLastError();
}

}  // namespace lam
