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
#include <algorithm>
#include <type_traits>

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
  // j in [0, ne), r in [0, ne - 1): (j - 1 + r) mod (ne - 1) needs one conditional subtract
  int v = j - 1 + r;
  if (v >= ne - 1) v -= ne - 1;
  return j == 0 ? 0 : 1 + v;
}

__device__ __forceinline__ float wave_sum(float v);
__device__ __forceinline__ float wave_max(float v);
__device__ __forceinline__ void wave_sync();
// Maximum over the workgroup of non-negative values.
__device__ __forceinline__ float block_max(float v, float *red) {
  v = wave_max(v);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float t = 0.0f;
  for (int i = 0; i < nw; ++i) t = fmaxf(t, red[i]);
  __syncthreads();
  return t;
}
__device__ __forceinline__ float block_sum(float v, float *red) {
  v = wave_sum(v);
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
#ifdef LAM_EIG_ROT32
  {
    const float b2 = b.x * b.x + b.y * b.y;
    if (b2 > 0.0f) {
      const float inv_b = rsqrtf(b2);
      const float tau = (aqq - app) * 0.5f * inv_b;
      const float t = 1.0f / (fabsf(tau) + sqrtf(1.0f + tau * tau));
      const float cc = rsqrtf(1.0f + t * t);
      R.c = cc;
      R.s = (tau >= 0.0f ? t : -t) * cc;
      R.er = b.x * inv_b;
      R.ei = b.y * inv_b;
    }
    return R;
  }
#endif
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

// Sizes of the Jacobi solver's LDS arrays.  jac_ne(n) = n + (n & 1): a zero
// dummy row and column pad odd n.  The input matrix is column-major with
// leading dimension jac_ne(n) (the solver's matrix buffers hold jac_ne(n)^2
// entries).  The eigenvectors V are column-major with leading dimension
// jac_ldv(n, nthr): its eigenvector threads give consecutive lanes of a pair
// consecutive rows and consecutive pairs consecutive columns, so with
// ldv = tpp (mod 32) (tpp: threads per pair) a wave's accesses tile the 32
// bank pairs.
__host__ __device__ constexpr int jac_ne(int n) { return n + (n & 1); }
__host__ __device__ inline void jac_vplan(int n, int nthr, int &vm, int &tpp) {
  const int h = jac_ne(n) >> 1, vcap = nthr > 64 ? nthr - 64 : nthr;
  vm = (h * n + vcap - 1) / vcap;
  while (h * ((n + vm - 1) / vm) > vcap) ++vm;
  tpp = (n + vm - 1) / vm;
}
__host__ __device__ inline int jac_ldv(int n, int nthr) {
  int vm, tpp;
  jac_vplan(n, nthr, vm, tpp);
  return n + (((tpp - n) % 32) + 32) % 32;
}

// LDS scratch of the Jacobi solver (carved by carve()).
struct JacobiLds {
  float2 *A1;  // second matrix buffer, jac_ne(n)^2
  Rot *rot;    // rotations [2][32], by round parity
  float *dn;   // rotated diagonals [2][64], by round parity
  float *red;  // per-wave partial sums
};

// 1/sqrt(x) in FP64 for x in [1e-36, 1e36]: the FP32 hardware estimate and one
// FP64 Newton step (about 46 correct bits), a shorter dependency than the FP64
// estimate.
__device__ __forceinline__ double rsq64f(double x) {
  const double y = (double)__builtin_amdgcn_rsqf((float)x);
  return y * (1.5 - 0.5 * (x * y) * y);
}

// Rotation that annihilates a_pq, as make_rotation (angles in FP64), in a form
// with a short dependent chain: with d = a_qq - a_pp, r = sqrt(d^2 + 4|b|^2),
// w = r + |d|, z = 1/sqrt(2 r w):
//   t = sgn(d) 2|b| / w  (= sgn(tau) / (|tau| + sqrt(1 + tau^2))),
//   c = 1/sqrt(1 + t^2) = w z,  s = t c = sgn(d) 2|b| z,
// and tb = t |b| = sgn(d) 2|b|^2 / w = sgn(d) 4 |b|^2 r z^2: the rotation leaves
// the diagonal a_pp - tb, a_qq + tb (the rotated diagonal in closed form).
// |a_pq|^2 below 1e-24 (a_pp^2 + a_qq^2) or 1e-36 is left in place (identity):
// such a rotation changes no FP32 entry.  Entries are assumed below 1e18.
__device__ __forceinline__ Rot make_rotation_t(float2 b, float app, float aqq, double &tb) {
  const double br = b.x, bi = b.y, d = (double)aqq - (double)app;
  const double b2 = br * br + bi * bi;
  const bool rot = b2 > 1e-36 && b2 > 1e-24 * ((double)app * app + (double)aqq * aqq);
  const double q = rot ? d * d + 4.0 * b2 : 1.0;
  const double yb = rsq64f(rot ? b2 : 1.0), bm = b2 * yb;
  const double r = q * rsq64f(q), w = r + fabs(d);
  const double z = rsq64f(2.0 * r * w);
  const double sg = d >= 0.0 ? 1.0 : -1.0;
  Rot R;
  R.c = rot ? (float)(w * z) : 1.0f;
  R.s = rot ? (float)(sg * 2.0 * bm * z) : 0.0f;
  R.er = rot ? (float)(br * yb) : 1.0f;
  R.ei = rot ? (float)(bi * yb) : 0.0f;
  tb = rot ? sg * 4.0 * b2 * r * z * z : 0.0;
  return R;
}

// Round-robin pairing in slot coordinates: pair k is slots (k, ne-1-k) in
// every round; between rounds the contents of slots 1..ne-1 move down one
// slot (slot 1 to slot ne-1), slot 0 stays.  After ne-1 rounds (a sweep) the
// slots are back in matrix order.
__device__ __forceinline__ int jac_sig(int s, int ne) { return s == 0 ? 0 : (s == 1 ? ne - 1 : s - 1); }
__device__ __forceinline__ int jac_sig_inv(int t, int ne) { return t == 0 ? 0 : (t == ne - 1 ? 1 : t + 1); }
__device__ __forceinline__ int jac_pair_of(int s, int ne) { return min(s, ne - 1 - s); }
// Storage of the off-diagonal entry (row slot a, column slot b): in the
// orientation with the row in the lower pair (within one pair: the lower
// slot; conj when stored transposed), as entry e = 2 [row is the pair's high
// slot] + [column is the high slot] of the 2x2 block (pair of row, pair of
// column), in plane e: offset e h^2 + kp h + kq.  A round's thread reads its
// block's entries from four planes and lanes of a wave hold neighbouring
// blocks, so the accesses of a wave are mostly consecutive words.
__device__ __forceinline__ int jac_slot_off(int a, int b, int ne, bool &conj) {
  const int h = ne >> 1;
  int pa = jac_pair_of(a, ne), pb = jac_pair_of(b, ne);
  const bool keep = pa < pb || (pa == pb && a < b);
  conj = !keep;
  if (!keep) {
    const int t = a; a = b; b = t;
    const int u = pa; pa = pb; pb = u;
  }
  const int e = 2 * (a >= h) + (b >= h);
  return (e * h + pa) * h + pb;
}
// Off-diagonal block (pair kp < pair kq) at list position pos of h pairs: the
// blocks holding the next round's pivots first (position k' holds the pivot
// of pair k' of the next round: (0,1), (k'-1,k'+1), (h-2,h-1)), then the
// others row by row.
__device__ __forceinline__ int2 jac_block(int pos, int h) {
  if (h == 2) return make_int2(0, 1);
  if (pos < h) return pos == 0 ? make_int2(0, 1) : (pos == h - 1 ? make_int2(h - 2, h - 1) : make_int2(pos - 1, pos + 1));
  int j = pos - h;
  for (int kp = 0; kp < h - 2; ++kp) {
    const int cnt = (h - 1 - kp) - (kp <= h - 3) - (kp == 0);
    if (j < cnt) return make_int2(kp, kp == 0 ? 3 + j : (j == 0 ? kp + 1 : kp + 2 + j));
    j -= cnt;
  }
  return make_int2(0, 1);
}

// Parallel cyclic Jacobi on the Hermitian n x n matrix A (column-major,
// leading dimension jac_ne(n), upper triangle read; n <= 64) in LDS; V
// (leading dimension jac_ldv(n, NTHR)) receives the eigenvectors (columns) and
// the diagonal of A the eigenvalues (unsorted); the rest of A is overwritten.
// Returns the number of sweeps.
//
// Each round rotates the ne/2 disjoint pairs (round-robin order) with one
// workgroup barrier.  The matrix lives in slot coordinates, double-buffered:
// a round reads its entries from one buffer and writes the rotated ones to
// the other at the slots they occupy in the next round, so every thread's
// addresses are fixed for the whole solve.  Only the off-diagonal 2x2 blocks
// of pair pairs are stored (one orientation); the annihilated pivots are
// exact zeros and the diagonal is carried in closed form (a_pp - t|a_pq|,
// a_qq + t|a_pq|) in a small table.  The thread whose block holds a pivot of
// the next round computes that rotation (angles in FP64) right after its
// update, so rotations never wait for a separate phase; threads of the last
// waves apply the round's rotations to the eigenvectors, V <- V J.
template <int NTHR>
__device__ __forceinline__ int jacobi_lds(float2 *A, float2 *V, int n, int max_sweeps, float tol,
                                          const JacobiLds &L) {
  constexpr int NW = NTHR / 64;
  constexpr int MAXB = (32 * 31 / 2 + NTHR - 1) / NTHR;  // off-diagonal blocks per thread at n = 64
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int ne = jac_ne(n), h = ne >> 1, nr = ne - 1, lda = ne, ldv = jac_ldv(n, NTHR);
  const int nb = h * (h - 1) / 2;
  // the buffers of the rounds: parity 0 reads A1, parity 1 reads A
  float2 *const B0 = L.A1, *const B1 = A;
  // input entry (r, c) from the upper triangle, zero on the dummy
  auto ain = [&](int r, int c) {
    if (r >= n || c >= n) return make_float2(0.f, 0.f);
    if (r == c) return make_float2(A[c * lda + c].x, 0.f);
    return r < c ? A[c * lda + r] : make_float2(A[r * lda + c].x, -A[r * lda + c].y);
  };
  // set-up: the off-diagonal entries into the planes of buffer 0, V = I, and
  // the convergence measure of the input: off = sum of |a_ij|^2 over i != j,
  // tot = off + sum a_ii^2
  {
    float off = 0.f, dg = 0.f;
    const float inv_ne = 1.0f / (float)ne;
    for (int i = tid; i < ne * ne; i += NTHR) {
      const int c = (int)(((float)i + 0.5f) * inv_ne), r = i - c * ne;
      const float2 a = ain(r, c);
      if (r < c) {
        bool cj;
        const int o = jac_slot_off(r, c, ne, cj);
        B0[o] = make_float2(a.x, cj ? -a.y : a.y);
      }
      if (r == c) dg += a.x * a.x;
      else off += a.x * a.x + a.y * a.y;
      if (r < n && c < n) V[c * ldv + r] = make_float2(r == c ? 1.0f : 0.0f, 0.0f);
    }
    off = wave_sum(off);
    dg = wave_sum(dg);
    if (lane == 0) {
      L.red[2 * w] = off;
      L.red[2 * w + 1] = dg;
    }
    __syncthreads();
    off = 0.f;
    dg = 0.f;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      off += L.red[2 * i];
      dg += L.red[2 * i + 1];
    }
    if (max_sweeps <= 0 || off == 0.0f || off <= tol * tol * (off + dg)) {
      __syncthreads();
      return 0;
    }
  }
  // rotations of round 0 (table 0)
  if (tid < h) {
    const int k = tid, kb = ne - 1 - k;
    double tbd;
    const float app = ain(k, k).x, aqq = ain(kb, kb).x;
    L.rot[k] = make_rotation_t(ain(k, kb), app, aqq, tbd);
    L.dn[k] = (float)((double)app - tbd);
    L.dn[kb] = (float)((double)aqq + tbd);
  }
  // this thread's blocks (slot j: list position tid + j NTHR): entry e = 2 i + j
  // is (row slot i of pair kp, column slot j of pair kq); read offset, write
  // offset in the other buffer, conj mask.  Threads without a block read
  // entry 0 and do not store.
  int rd[MAXB][4], wr[MAXB][4], bkp[MAXB], bkq[MAXB];
  unsigned cjm[MAXB];
  bool bval[MAXB];
#pragma unroll
  for (int j = 0; j < MAXB; ++j) {
    const int pos = tid + j * NTHR;
    bval[j] = pos < nb;
    bkp[j] = bkq[j] = 0;
    cjm[j] = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) rd[j][e] = wr[j][e] = 0;
    if (bval[j]) {
      const int2 B = jac_block(pos, h);
      bkp[j] = B.x;
      bkq[j] = B.y;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int a = (e >> 1) ? ne - 1 - B.x : B.x, b = (e & 1) ? ne - 1 - B.y : B.y;
        bool cj;
        rd[j][e] = (e * h + B.x) * h + B.y;
        wr[j][e] = jac_slot_off(jac_sig(a, ne), jac_sig(b, ne), ne, cj);
        cjm[j] |= cj ? (0x80000000u >> e) : 0u;
      }
    }
  }
  // pivot tasks (wave 0): the next round's rotation of pair tk from entry te
  // of block 0 (te < 0: the pivot is the zero just annihilated, n <= 2), its
  // diagonal from slots ta, tq of this round's table.  One task per thread
  // k < h; with h = 2 thread 0 has both.
  int tk[2] = {0, 0}, te[2] = {-1, -1}, ta[2] = {0, 0}, tq[2] = {0, 0};
  bool tval[2] = {false, false};
  {
    const int first = h == 2 ? 0 : tid, count = h == 2 ? (tid == 0 ? 2 : 0) : (tid < h ? 1 : 0);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      if (t < count) {
        const int kn = first + t, a0 = jac_sig_inv(kn, ne), b0 = jac_sig_inv(ne - 1 - kn, ne);
        int e = -1;
        if (h > 1) {
          for (int ee = 0; ee < 4; ++ee) {
            const int a = (ee >> 1) ? ne - 1 - bkp[0] : bkp[0], b = (ee & 1) ? ne - 1 - bkq[0] : bkq[0];
            if ((a == a0 && b == b0) || (a == b0 && b == a0)) e = ee;
          }
        }
        tk[t] = kn;
        te[t] = e;
        ta[t] = a0;
        tq[t] = b0;
        tval[t] = true;
      }
    }
  }
  // the zero that replaces this round's pivot of pair tid in the next buffer
  int zoff = 0;
  if (tid < h) {
    bool cj;
    zoff = jac_slot_off(jac_sig(tid, ne), jac_sig(ne - 1 - tid, ne), ne, cj);
  }
  // eigenvector items: (pair vk, rows vj, vj + tpp, ...) on the last nvt threads
  // (wave 0 keeps the rotations), at most vm rows each; consecutive lanes take
  // consecutive rows, so an access spreads over the LDS banks
  int vm, tpp;
  jac_vplan(n, NTHR, vm, tpp);
  const int nvt = h * tpp, vt = tid - (NTHR - nvt);
  const bool vthread = vt >= 0;
  const int vk = vthread ? vt / tpp : 0, vj = vthread ? vt - vk * tpp : 0;
  const int vrows = vthread ? (n - vj + tpp - 1) / tpp : 0;
  int vcp = vk, vcq = ne - 1 - vk;  // matrix columns of slots vk, ne-1-vk in round 0
  int vop = vcp * ldv + vj, voq = vcq * ldv + vj;  // their element offsets (row vj)
  int vro[4];  // offsets of the first four rows (the last valid row repeated)
#pragma unroll
  for (int u = 0; u < 4; ++u) vro[u] = min(u, max(vrows - 1, 0)) * tpp;
  // per-wave roles (uniform): block slots, pivot tasks, eigenvector items
  const int wlo = w * 64;
  const bool wblk0 = wlo < nb, wblk1 = MAXB > 1 && wlo + NTHR < nb;
  const bool wtask = w == 0;
  const bool wv = wlo + 64 > NTHR - nvt;
  __syncthreads();

  // one block: A <- J_P^H A J_Q on its 2x2 entries, stored rotated and moved
  auto block = [&](int j, const float2 *Sc, float2 *Sn, const Rot *rc, float2 (&x)[4], bool last,
                   float &offp) {
    const Rot RP = rc[bkp[j]], RQ = rc[bkq[j]];
    rotate_block(RP, RQ, x[0], x[1], x[2], x[3]);
    const unsigned m = cjm[j];
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e].y = __uint_as_float(__float_as_uint(x[e].y) ^ ((m << e) & 0x80000000u));
    if (bval[j]) {
#pragma unroll
      for (int e = 0; e < 4; ++e) Sn[wr[j][e]] = x[e];
    }
    if (last) {  // uniform: the last round of a sweep
      const float m2 = x[0].x * x[0].x + x[0].y * x[0].y + x[1].x * x[1].x + x[1].y * x[1].y +
                       x[2].x * x[2].x + x[2].y * x[2].y + x[3].x * x[3].x + x[3].y * x[3].y;
      offp += bval[j] ? 2.0f * m2 : 0.0f;
    }
  };
  // V <- V J on up to 4 rows (stride tpp) of this thread's item
  auto vrot = [&](const Rot &R, float2 *cp, float2 *cq, const float2 (&vp)[4], const float2 (&vq)[4], int rows) {
    const float2 eb = make_float2(R.er, -R.ei);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float2 y = cmul(eb, vq[u]);
      if (u < rows) {
        cp[u * tpp] = make_float2(R.c * vp[u].x - R.s * y.x, R.c * vp[u].y - R.s * y.y);
        cq[u * tpp] = make_float2(R.s * vp[u].x + R.c * y.x, R.s * vp[u].y + R.c * y.y);
      }
    }
  };

  // One round, straight-line per wave role: every load first (the next
  // pivots' diagonal, block 0, the first eigenvector rows), then block 0 and
  // the next rotations (wave 0), further blocks, the eigenvector update.
  auto round = [&](auto parity, auto task, bool last) {
    constexpr int PAR = decltype(parity)::value;
    constexpr bool TASK = decltype(task)::value;
    const float2 *Sc = PAR ? B1 : B0;
    float2 *Sn = PAR ? B0 : B1;
    const Rot *rc = L.rot + PAR * 32;
    Rot *rn = L.rot + (1 - PAR) * 32;
    const float *dc = L.dn + PAR * 64;
    float *dnx = L.dn + (1 - PAR) * 64;
    float tpa[2] = {0.f, 0.f}, tpq[2] = {0.f, 0.f};
    if constexpr (TASK) {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        tpa[t] = dc[ta[t]];
        tpq[t] = dc[tq[t]];
      }
    }
    float2 x[4];
    if (wblk0) {
#pragma unroll
      for (int e = 0; e < 4; ++e) x[e] = Sc[rd[0][e]];
    }
    const bool vact = vthread && vcp < n && vcq < n;
    float2 *cp = V + vop, *cq = V + voq;
    float2 vp[4], vq[4];
    Rot RV;
    if (wv) {
      RV = rc[vk];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        vp[u] = cp[vro[u]];
        vq[u] = cq[vro[u]];
      }
    }
    float offp = 0.f;
    if (wblk0) {
      block(0, Sc, Sn, rc, x, last, offp);
      if constexpr (TASK) {
        // the next round's rotations of the pivots this thread holds
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          if (t == 0 || h == 2) {
            const int e = te[t];
            const float2 pv = e == 0 ? x[0] : (e == 1 ? x[1] : (e == 2 ? x[2] : (e == 3 ? x[3] : make_float2(0.f, 0.f))));
            double tbd;
            const Rot R = make_rotation_t(pv, tpa[t], tpq[t], tbd);
            if (tval[t]) {
              rn[tk[t]] = R;
              dnx[tk[t]] = (float)((double)tpa[t] - tbd);
              dnx[ne - 1 - tk[t]] = (float)((double)tpq[t] + tbd);
            }
          }
        }
      }
    } else if constexpr (TASK) {
      // n <= 2: no blocks; the next rotation is the identity on the rotated diagonal
      if (tval[0]) {
        rn[tk[0]] = Rot{1.0f, 0.0f, 1.0f, 0.0f};
        dnx[tk[0]] = tpa[0];
        dnx[ne - 1 - tk[0]] = tpq[0];
      }
    }
    if (tid < h) Sn[zoff] = make_float2(0.f, 0.f);
    if (wblk1) {
#pragma unroll
      for (int j = 1; j < MAXB; ++j) {
        if (wlo + j * NTHR < nb) {
          float2 y[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) y[e] = Sc[rd[j][e]];
          block(j, Sc, Sn, rc, y, last, offp);
        }
      }
    }
    if (wv) {
      if (vact) vrot(RV, cp, cq, vp, vq, vrows);
      for (int r0 = 4; r0 < vrows; r0 += 4) {
        float2 wp[4], wq[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int o = min(r0 + u, vrows - 1) * tpp;
          wp[u] = cp[o];
          wq[u] = cq[o];
        }
        if (vact) vrot(RV, cp + r0 * tpp, cq + r0 * tpp, wp, wq, vrows - r0);
      }
    }
    // next round: slot contents move down one slot, so the matrix column of a
    // slot (other than 0) advances by one, ne - 1 wrapping to 1
    if (vk != 0) {
      const bool wp = vcp + 1 == ne;
      vcp = wp ? 1 : vcp + 1;
      vop = wp ? ldv + vj : vop + ldv;
    }
    {
      const bool wq = vcq + 1 == ne;
      vcq = wq ? 1 : vcq + 1;
      voq = wq ? ldv + vj : voq + ldv;
    }
    if (last) {
      offp = wave_sum(offp);
      if (lane == 0) L.red[w] = offp;
    }
    __syncthreads();
  };
  auto round_any = [&](auto parity, bool last) {
    if (wtask) round(parity, std::true_type{}, last);
    else round(parity, std::false_type{}, last);
  };
  // converged after the last round of a sweep (table par holds its diagonal)
  auto converged = [&](int par) {
    float offs = 0.f;
#pragma unroll
    for (int i = 0; i < NW; ++i) offs += L.red[i];
    const float d = lane < ne ? L.dn[par * 64 + lane] : 0.f;
    const float tot = offs + wave_sum(d * d);
    return offs == 0.0f || offs <= tol * tol * tot;
  };
  int sweep = 0, r = 0, par_last = 0;
  for (;;) {
    round_any(std::integral_constant<int, 0>{}, r == nr - 1);
    if (++r == nr) {
      r = 0;
      par_last = 0;
      if (++sweep >= max_sweeps || converged(0)) break;
    }
    round_any(std::integral_constant<int, 1>{}, r == nr - 1);
    if (++r == nr) {
      r = 0;
      par_last = 1;
      if (++sweep >= max_sweeps || converged(1)) break;
    }
  }
  // eigenvalues onto the diagonal of A: matrix index i sat in slot
  // jac_sig_inv(i) of the last round
  if (tid < n) A[tid * lda + tid] = make_float2(L.dn[par_last * 64 + jac_sig_inv(tid, ne)], 0.f);
  __syncthreads();
  return sweep;
}

// Ascending order of the diagonal of A: perm[rank] = column, ev[rank] = value.
__device__ __forceinline__ void sort_eigen(const float2 *A, int n, int *perm, float *ev) {
  const int j = threadIdx.x, lda = jac_ne(n);
  if (j < n) {
    const float lj = A[j * lda + j].x;
    int rank = 0;
    for (int k = 0; k < n; ++k) {
      const float lk = A[k * lda + k].x;
      rank += (lk < lj) || (lk == lj && k < j);
    }
    perm[rank] = j;
    ev[rank] = lj;
  }
  __syncthreads();
}

// Linear-interpolated percentile (pipeline/common.hpp:461-478), FP32 like the reference.
__device__ __forceinline__ float pct_linear(const float *v, int n, float q) {
  if (n <= 0) return 0.0f;
  if (n == 1) return v[0];
  const float pos = fminf(1.0f, fmaxf(0.0f, q)) * (float)(n - 1);
  const int lo = (int)floorf(pos), hi = (int)ceilf(pos);
  if (lo == hi) return v[lo];
  return v[lo] + (pos - (float)lo) * (v[hi] - v[lo]);
}

struct EigSmem {
  float2 *A, *V, *W, *G;           // A: leading dimension jac_ne(n); V: ldv
  int ldv;
  float2 *A0, *Y, *hv, *t1, *t2;  // warm start: copy of A, proposal, reflectors, scratch
  float *ev, *d, *f, *red, *htau, *res;
  int *perm;
  int2 *rpq;                      // small (warm-start) Jacobi: pairs (shares pivo)
  Rot *rot;                       // small (warm-start) Jacobi: rotations (shares rotw)
  Rot *rot2;                      // Jacobi solver: rotations [2][32]
  int2 *rpq2;                     // Jacobi solver: rotated diagonals [2][64] (as floats)
  int *kdet, *mark;
  __device__ JacobiLds jl() const { return JacobiLds{W, rot2, reinterpret_cast<float *>(rpq2), red}; }
};

// ---- plain solve without eigen outputs: Householder reduction to a real
// tridiagonal, all eigenvalues by multisection on Sturm counts, the removed
// modes' eigenvectors by inverse iteration and the reflectors ----
// The eigenvectors of a problem are wanted only for its removed modes (and
// the next block's warm start), so the O(n^3) work of the Jacobi solver's
// accumulated rotations is replaced by one reduction (about 16/3 n^3 real
// flops, one barrier per column) and O(n^2) per vector.
constexpr int kTriK = 8;  // eigenvectors per problem (more: the Jacobi solver)

// LDS of the tridiagonal path inside the W region (nothing else there is
// live until the weights are loaded): the reduction's partial sums, then the
// elimination rows and the tridiagonal's eigenvectors; the tables after them.
struct TriLds {
  float2 *part;          // [2][waves][64] partial products A22 u, then [2][64] next columns
  float *fac;            // [n][4][kTriK] elimination rows: 1/pivot, u1, u2, rhs
  float *z;              // [n][kTriK] eigenvectors of the real tridiagonal
  float *ta, *tb, *hc;   // [64] diagonal, off-diagonal (>= 0), reflector scales
  float2 *ph;            // [64] off-diagonal phases, then the diagonal similarity
  int *flag;             // [1] vectors finite
};
__host__ __device__ inline int tri_w_base(int n) { return 20 * n > 640 ? 20 * n : 640; }
__host__ __device__ inline int tri_w_elems(int n) { return tri_w_base(n) + 256; }

// LDS layout.  The warm-start buffers exist only in the warm kernel.  W
// (loaded after the eigen solve) also holds the solver's second matrix buffer
// and the warm start's copy of A (used before any plain solve).  In the plain
// kernel G (formed after the sorted basis has been copied out of V) reuses V:
// 34.0 KB at n = 36, B = 16, so three eigen workgroups leave room on a CU for
// a voltage workgroup of the neighbouring blocks (55 KB) instead of blocking it.
__host__ __device__ inline int eig_w_elems(int n, int B, bool warm, int nthr) {
  const int m = jac_ne(n) * jac_ne(n), mv = jac_ne(n) * jac_ldv(n, nthr);
  const bool g_in_v = !warm && B * n <= mv;  // G (B x n) fits in V
  int wb = (g_in_v ? 2 : 3) * B * n;
  if (!warm && tri_w_elems(n) > wb) wb = tri_w_elems(n);  // the tridiagonal path's scratch
  return wb > m ? wb : m;
}

__device__ __forceinline__ EigSmem carve(char *base, int n, int B, bool warm, int nw) {
  EigSmem s;
  const int m = jac_ne(n) * jac_ne(n);
  s.ldv = jac_ldv(n, nw * 64);
  const int mv = jac_ne(n) * s.ldv;
  s.A = reinterpret_cast<float2 *>(base);
  s.V = s.A + m;
  s.W = s.V + mv;
  const bool g_in_v = !warm && B * n <= mv;
  s.G = g_in_v ? s.V : s.W + 2 * B * n;
  float2 *q = s.W + eig_w_elems(n, B, warm, nw * 64);
  s.A0 = s.Y = s.hv = s.t1 = s.t2 = nullptr;
  if (warm) {
    s.A0 = s.W;
    s.Y = q;
    s.hv = s.Y + kMaxWarm * n;
    s.t1 = s.hv + kMaxWarm * n;
    s.t2 = s.t1 + 64;
    q = s.t2 + 64;
  }
  float *fp = reinterpret_cast<float *>(q);
  s.ev = fp;
  s.d = fp + 64;
  s.f = fp + 128;
  s.red = fp + 192;
  s.htau = fp + 208;
  s.res = fp + 216;  // kMaxWarm * 64 (warm only)
  s.perm = reinterpret_cast<int *>(s.res + (warm ? kMaxWarm * 64 : 0));
  s.mark = s.perm + 64;
  // 16-byte alignment by offset arithmetic on the LDS base (an integer cast of
  // the pointer would turn every table access into a flat access)
  const long moff = (reinterpret_cast<char *>(s.mark + 64) - base + 15) & ~15L;
  s.rot2 = reinterpret_cast<Rot *>(base + moff);
  s.rpq2 = reinterpret_cast<int2 *>(s.rot2 + 64);
  s.kdet = reinterpret_cast<int *>(s.rpq2 + 64);
  // the warm start's small Jacobi (wave 0, never beside the solver) reuses them
  s.rot = s.rot2;
  s.rpq = s.rpq2;
  (void)nw;
  return s;
}

size_t eig_smem_bytes(int n, int B, bool warm, int nw);
// LDS of a workgroup that takes problems of any size up to max_n (the
// eigenvector stride, hence the size, is not monotone in n)
size_t eig_smem_bytes_upto(int max_n, int B, bool warm, int nw) {
  size_t m = 0;
  for (int n = 1; n <= max_n; ++n) m = std::max(m, eig_smem_bytes(n, B, warm, nw));
  return m;
}

size_t eig_smem_bytes(int n, int B, bool warm, int nw) {
  size_t f2 = (size_t)jac_ne(n) * (jac_ne(n) + jac_ldv(n, nw * 64)) + eig_w_elems(n, B, warm, nw * 64);
  if (warm) f2 += 2 * kMaxWarm * n + 128;
  return sizeof(float2) * f2 + sizeof(float) * (216 + (warm ? kMaxWarm * 64 : 0)) + sizeof(int) * 128 +
         15 + sizeof(Rot) * 64 + sizeof(int2) * 64 + sizeof(int) * 4 + 0 * nw;
}

// Sum over the 64 lanes of a wavefront with DPP lane moves (row_shr 1/2/4/8
// leaves each 16-lane row's total in its lane 15; row_bcast 15 / 31 carry the
// row totals up to lane 63), broadcast back with readlane.  All lanes active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                                               ROW_MASK, 0xf, false));
}
// Lane l of v (l uniform).
__device__ __forceinline__ float rlf(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ float2 rl2(float2 v, int l) { return make_float2(rlf(v.x, l), rlf(v.y, l)); }
// Sum over the 64 lanes in FP64 (the DPP pattern of wave_sum on both halves).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move_f64(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_sum_f64(double v) {
  v += dpp_move_f64<0x111, 0xf>(v);
  v += dpp_move_f64<0x112, 0xf>(v);
  v += dpp_move_f64<0x114, 0xf>(v);
  v += dpp_move_f64<0x118, 0xf>(v);
  v += dpp_move_f64<0x142, 0xa>(v);
  v += dpp_move_f64<0x143, 0xc>(v);
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_move<0x111, 0xf>(v);  // row_shr:1
  v += dpp_move<0x112, 0xf>(v);  // row_shr:2
  v += dpp_move<0x114, 0xf>(v);  // row_shr:4
  v += dpp_move<0x118, 0xf>(v);  // row_shr:8
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}
// Maximum over the 64 lanes of non-negative values (lanes outside a row read
// 0 in the DPP moves, neutral for non-negative inputs).
__device__ __forceinline__ float wave_max(float v) {
  v = fmaxf(v, dpp_move<0x111, 0xf>(v));
  v = fmaxf(v, dpp_move<0x112, 0xf>(v));
  v = fmaxf(v, dpp_move<0x114, 0xf>(v));
  v = fmaxf(v, dpp_move<0x118, 0xf>(v));
  v = fmaxf(v, dpp_move<0x142, 0xa>(v));
  v = fmaxf(v, dpp_move<0x143, 0xc>(v));
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

// Number of modes to remove: fixed K, or the detection rule of
// detectSignalEigenmodeCountsKernel (spatial.cuh:900-950) on ascending ev.
__device__ __forceinline__ int select_k(const float *ev, int n, int fixed, int detect, float delta, int *kdet) {
  if (threadIdx.x == 0) {
    int k = fixed;
    if (detect) {
      const float p20 = pct_linear(ev, n, 0.2f), p50 = pct_linear(ev, n, 0.5f),
                  p80 = pct_linear(ev, n, 0.8f);
      const float thr = p50 + delta * ((p80 - p20) / (2.0f * 0.8416f));
      k = 0;
      for (int i = 0; i < n; ++i) k += ev[i] > thr;
    }
    kdet[0] = max(0, min(n, k));
  }
  __syncthreads();
  return kdet[0];
}

// Wave-level ordering of LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Z[k][i] = sum_j A(i, j) Y[k][j] for KW vectors (A column-major, stride ld).
__device__ __forceinline__ void apply_cols(const float2 *A, int ld, const float2 *Y, float2 *Z, int n,
                                           int KW) {
  for (int e = threadIdx.x; e < KW * n; e += blockDim.x) {
    const int k = e / n, i = e - k * n;
    // four independent partial sums: a short dependent chain per thread
    float2 z[4] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    int j = 0;
    for (; j + 4 <= n; j += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float2 t = cmul(A[(j + u) * ld + i], Y[k * n + j + u]);
        z[u].x += t.x;
        z[u].y += t.y;
      }
    }
    for (; j < n; ++j) {
      const float2 t = cmul(A[j * ld + i], Y[k * n + j]);
      z[0].x += t.x;
      z[0].y += t.y;
    }
    Z[e] = make_float2((z[0].x + z[1].x) + (z[2].x + z[3].x), (z[0].y + z[1].y) + (z[2].y + z[3].y));
  }
}

// Classical Gram-Schmidt applied twice ("twice is enough") on KW vectors
// Y[k][0..n) by wave 0 (lane = row, n <= 64): the projections of vector k on
// all earlier vectors are independent wave sums, issued together.  A vector
// that vanishes is replaced by a unit vector before orthogonalising.
__device__ __forceinline__ void mgs_wave(float2 *Y, int n, int KW) {
  const int i = threadIdx.x;
  float2 y[kMaxWarm];
#pragma unroll
  for (int k = 0; k < kMaxWarm; ++k) y[k] = (k < KW && i < n) ? Y[k * n + i] : make_float2(0.f, 0.f);
#pragma unroll
  for (int k = 0; k < kMaxWarm; ++k) {
    if (k < KW) {
      for (int attempt = 0; attempt < 2; ++attempt) {
        const float n0 = wave_sum(y[k].x * y[k].x + y[k].y * y[k].y);
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
          float gr[kMaxWarm], gi[kMaxWarm];
#pragma unroll
          for (int j = 0; j < k; ++j) {  // conj(y_j) . y_k
            gr[j] = wave_sum(y[j].x * y[k].x + y[j].y * y[k].y);
            gi[j] = wave_sum(y[j].x * y[k].y - y[j].y * y[k].x);
          }
#pragma unroll
          for (int j = 0; j < k; ++j) {
            y[k].x -= gr[j] * y[j].x - gi[j] * y[j].y;
            y[k].y -= gr[j] * y[j].y + gi[j] * y[j].x;
          }
        }
        const float nrm2 = wave_sum(y[k].x * y[k].x + y[k].y * y[k].y);
        if (nrm2 > 1e-12f * n0 && nrm2 > 1e-30f) {
          const float inv = rsqrtf(nrm2);
          y[k].x *= inv;
          y[k].y *= inv;
          break;
        }
        // (nearly) dependent: restart from a unit vector
        y[k] = make_float2(i == (k * 5 + 3 + attempt * 7) % n ? 1.0f : 0.0f, 0.0f);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kMaxWarm; ++k)
    if (k < KW && i < n) Y[k * n + i] = y[k];
}

// Cyclic Jacobi on the small Hermitian KW x KW matrix B (row-major, KW even)
// by wave 0 alone: lane (a, b) owns B(a, b) and Vb(a, b); rotation angles in FP64.
__device__ __forceinline__ void small_jacobi_wave(float2 *B, float2 *Vb, int KW, int max_sweeps,
                                                  float tol, EigSmem &S) {
  const int l = threadIdx.x, a = l >> 3, b = l & 7;
  const bool act = a < KW && b < KW;
  if (act) {
    Vb[a * KW + b] = make_float2(a == b ? 1.0f : 0.0f, 0.0f);
    if (a == b) B[a * KW + b].y = 0.0f;
  }
  wave_sync();
  for (int sw = 0; sw < max_sweeps; ++sw) {
    float2 x = act ? B[a * KW + b] : make_float2(0.f, 0.f);
    const float m = x.x * x.x + x.y * x.y;
    const float off = wave_sum(a != b ? m : 0.0f), tot = wave_sum(m);
    if (off == 0.0f || off <= tol * tol * tot) break;
    for (int r = 0; r < KW - 1; ++r) {
      if (l < KW / 2) {
        int p = rr_pos(l, r, KW), q = rr_pos(KW - 1 - l, r, KW);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double tbd;
        S.rot[l] = make_rotation_t(B[p * KW + q], B[p * KW + p].x, B[q * KW + q].x, tbd);
        S.rpq[l] = make_int2(p, q);
        S.mark[p] = 2 * l;
        S.mark[q] = 2 * l + 1;
      }
      wave_sync();
      float2 nb = make_float2(0.f, 0.f), nv = make_float2(0.f, 0.f);
      if (act) {
        const int sa = S.mark[a], sb = S.mark[b];
        const int2 Pa = S.rpq[sa >> 1], Pb = S.rpq[sb >> 1];
        const Rot Ra = S.rot[sa >> 1], Rb = S.rot[sb >> 1];
        // B(i, j) = B[i * KW + j]; rotate_block expects X(r, c) with rows of pair a
        float2 x00 = B[Pa.x * KW + Pb.x], x01 = B[Pa.x * KW + Pb.y];
        float2 x10 = B[Pa.y * KW + Pb.x], x11 = B[Pa.y * KW + Pb.y];
        rotate_block(Ra, Rb, x00, x01, x10, x11);
        const int ra = sa & 1, rb = sb & 1;
        nb = ra ? (rb ? x11 : x10) : (rb ? x01 : x00);
        if ((sa >> 1) == (sb >> 1)) {  // the annihilated pair
          if (ra != rb) nb = make_float2(0.f, 0.f);
          else nb.y = 0.f;
        }
        const float2 vp = Vb[a * KW + Pb.x], vq = Vb[a * KW + Pb.y];
        const float2 y = cmul(make_float2(Rb.er, -Rb.ei), vq);
        nv = rb ? make_float2(Rb.s * vp.x + Rb.c * y.x, Rb.s * vp.y + Rb.c * y.y)
                : make_float2(Rb.c * vp.x - Rb.s * y.x, Rb.c * vp.y - Rb.s * y.y);
      }
      wave_sync();
      if (act) {
        B[a * KW + b] = nb;
        Vb[a * KW + b] = nv;
      }
      wave_sync();
    }
  }
}

// Warm-start eigen solve of the whitened matrix A (stride ld, not modified):
//   Y <- proposal; `steps` times Y <- orth(A Y); Rayleigh-Ritz on Y^H A Y
//   (small Jacobi); Ritz vectors U = Y Vb; an orthonormal complement of U from
//   its Householder QR with Rayleigh quotients as the remaining eigenvalues.
// Fills S.V (columns: complement, then Ritz vectors by ascending Ritz value),
// S.ev / S.perm (ascending) and returns the accepted flag of the leakage test.
// With full == false (no eigen outputs requested) no complement is formed:
// its columns of S.V are left unset, its eigenvalues all read as their mean
// (trace(A) minus the Ritz values, over n - KW), and the complement's largest
// eigenvalue for the check below is estimated by three power steps on
// (I - U U^H) A (I - U U^H).
//   sum_{top K} |A u_k - theta_k u_k|^2 / theta_k * (theta_k - mu) / theta_k <= tol * mu,
// mu = mean of the remaining eigenvalues: the interference a mode leaks through
// the null is its angle error squared times its power above the noise floor.
// A complement direction ranked among the K largest fails the test when its
// Rayleigh quotient exceeds 1.5 mu (an interferer the warm subspace missed).
__device__ __forceinline__ bool warm_solve(const float2 *A, int ld, int n, int KW, int K, int steps,
                                           float tol, EigSmem &S, bool full = true) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  float2 *Y = S.Y, *Z = S.hv;
  for (int st = 0; st < steps; ++st) {
    apply_cols(A, ld, Y, Z, n, KW);
    __syncthreads();
    for (int e = tid; e < KW * n; e += nthr) Y[e] = Z[e];
    __syncthreads();
    if (tid < 64) mgs_wave(Y, n, KW);
    __syncthreads();
  }
  if (steps == 0) {
    if (tid < 64) mgs_wave(Y, n, KW);
    __syncthreads();
  }
  // Rayleigh-Ritz: B = Y^H A Y
  apply_cols(A, ld, Y, Z, n, KW);
  __syncthreads();
  float2 *Bm = S.t1, *Vb = S.t2;
  for (int e = tid; e < KW * KW; e += nthr) {
    const int ra = e / KW, cb = e - ra * KW;
    float2 g[4] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    for (int i0 = 0; i0 < n; i0 += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = min(i0 + u, n - 1);
        const float2 t = cmulc(Z[cb * n + i], Y[ra * n + i]);  // conj(y_a) z_b
        const float on = i0 + u < n ? 1.0f : 0.0f;
        g[u].x += on * t.x;
        g[u].y += on * t.y;
      }
    }
    Bm[ra * KW + cb] = make_float2((g[0].x + g[1].x) + (g[2].x + g[3].x), (g[0].y + g[1].y) + (g[2].y + g[3].y));
  }
  __syncthreads();
  if (tid < 64) small_jacobi_wave(Bm, Vb, KW, 12, 1e-7f, S);
  __syncthreads();
  // Ritz values ascending: rank of theta_b
  int *rk = S.mark;  // reuse: rank -> Ritz index
  if (tid < KW) {
    const float tb = Bm[tid * KW + tid].x;
    int rank = 0;
    for (int c2 = 0; c2 < KW; ++c2) {
      const float tc = Bm[c2 * KW + c2].x;
      rank += (tc < tb) || (tc == tb && c2 < tid);
    }
    rk[rank] = tid;
  }
  __syncthreads();
  // Ritz vectors into the last KW columns of V and residuals r = Z vb - theta u
  const int m = n - KW;
  for (int e = tid; e < KW * n; e += nthr) {
    const int k = e / n, i = e - k * n;
    const int bcol = rk[k];
    const float th = Bm[bcol * KW + bcol].x;
    float2 u = make_float2(0.f, 0.f), z = make_float2(0.f, 0.f);
    for (int j = 0; j < KW; ++j) {
      const float2 vb = Vb[j * KW + bcol];
      const float2 t1 = cmul(Y[j * n + i], vb), t2 = cmul(Z[j * n + i], vb);
      u.x += t1.x;
      u.y += t1.y;
      z.x += t2.x;
      z.y += t2.y;
    }
    S.V[(m + k) * S.ldv + i] = u;
    z.x -= th * u.x;
    z.y -= th * u.y;
    S.res[k * 64 + i] = z.x * z.x + z.y * z.y;
  }
  __syncthreads();
  float *evu = S.f;  // unsorted eigenvalues by column
  float cmax = 0.0f;  // complement's largest eigenvalue (estimate, full == false)
  if (full) {
    // complement: Householder QR of U (wave 0) and the compact form
    // H_0 ... H_{KW-1} = I - Vh T Vh^H (T upper triangular, built in the same pass);
    // V[:, 0..m) = (I - Vh T Vh^H)[:, KW..n)
    float2 *Tm = S.t2;  // KW x KW, row-major (Vb is no longer needed)
    if (tid < 64) {
      const int i = tid;
      for (int k = 0; k < KW; ++k)
        if (i < n) Y[k * n + i] = S.V[(m + k) * S.ldv + i];
      for (int e = i; e < KW * KW; e += 64) Tm[e] = make_float2(0.f, 0.f);
      wave_sync();
      float2 vk[kMaxWarm];
#pragma unroll
      for (int k = 0; k < kMaxWarm; ++k) {
        if (k >= KW) break;
        const float2 x = (i < n && i >= k) ? Y[k * n + i] : make_float2(0.f, 0.f);
        const float nrm2 = wave_sum(x.x * x.x + x.y * x.y);
        const float xkr = __shfl(x.x, k, 64), xki = __shfl(x.y, k, 64);
        const float axk = sqrtf(xkr * xkr + xki * xki), nrm = sqrtf(nrm2);
        const float2 alpha = axk > 0.0f ? make_float2(-xkr / axk * nrm, -xki / axk * nrm)
                                        : make_float2(-nrm, 0.0f);
        float2 v = x;
        if (i == k) {
          v.x -= alpha.x;
          v.y -= alpha.y;
        }
        const float vn2 = wave_sum(v.x * v.x + v.y * v.y);
        const float tau = (nrm2 > 1e-30f && vn2 > 1e-30f) ? 2.0f / vn2 : 0.0f;
        vk[k] = v;
        if (i < n) Z[k * n + i] = v;  // reflector k
        // T(0:k, k) = -tau T(0:k, 0:k) (Vh(:, 0:k)^H v_k), T(k, k) = tau
        float zr[kMaxWarm], zi[kMaxWarm];
#pragma unroll
        for (int j = 0; j < k; ++j) {
          zr[j] = wave_sum(vk[j].x * v.x + vk[j].y * v.y);
          zi[j] = wave_sum(vk[j].x * v.y - vk[j].y * v.x);
        }
        if (i < k) {
          float2 acc = make_float2(0.f, 0.f);
#pragma unroll
          for (int j = 0; j < k; ++j) {
            const float2 t = cmul(Tm[i * KW + j], make_float2(zr[j], zi[j]));
            acc.x += t.x;
            acc.y += t.y;
          }
          Tm[i * KW + k] = make_float2(-tau * acc.x, -tau * acc.y);
        }
        if (i == k) Tm[k * KW + k] = make_float2(tau, 0.f);
        wave_sync();
        // apply H_k to the remaining columns of U
#pragma unroll
        for (int j = k + 1; j < kMaxWarm; ++j) {
          if (j >= KW) break;
          float2 y = i < n ? Y[j * n + i] : make_float2(0.f, 0.f);
          const float gr = wave_sum(v.x * y.x + v.y * y.y), gi = wave_sum(v.x * y.y - v.y * y.x);
          y.x -= tau * (gr * v.x - gi * v.y);
          y.y -= tau * (gr * v.y + gi * v.x);
          if (i < n) Y[j * n + i] = y;
        }
        wave_sync();
      }
    }
    __syncthreads();
    // Sm(a, c) = sum_b T(a, b) conj(Vh(KW + c, b)), then Q2(i, c) = delta - sum_a Vh(i, a) Sm(a, c)
    float2 *Sm = Y;  // KW x m, column c at Sm + c * KW
    for (int e = tid; e < KW * m; e += nthr) {
      const int c2 = e / KW, a2 = e - c2 * KW;
      float2 acc = make_float2(0.f, 0.f);
      for (int b2 = a2; b2 < KW; ++b2) {
        const float2 t = cmulc(Tm[a2 * KW + b2], Z[b2 * n + KW + c2]);
        acc.x += t.x;
        acc.y += t.y;
      }
      Sm[c2 * KW + a2] = acc;
    }
    __syncthreads();
    for (int e = tid; e < m * n; e += nthr) {
      const int c2 = e / n, i = e - c2 * n;
      float2 q = make_float2(i == KW + c2 ? 1.0f : 0.0f, 0.0f);
      for (int a2 = 0; a2 < KW; ++a2) {
        const float2 t = cmul(Z[a2 * n + i], Sm[c2 * KW + a2]);
        q.x -= t.x;
        q.y -= t.y;
      }
      S.A0[e] = q;
    }
    __syncthreads();
    for (int e = tid; e < m * n; e += nthr) S.V[(e / n) * S.ldv + e % n] = S.A0[e];
    __syncthreads();
    // Rayleigh quotients of the complement: lambda_j = q_j^H A q_j
    for (int e = tid; e < m * n; e += nthr) {  // W = A Q2 into A0
      const int j = e / n, i = e - j * n;
      float2 w[4] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
      for (int c0 = 0; c0 < n; c0 += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int c2 = min(c0 + u, n - 1);
          const float2 t = cmul(A[c2 * ld + i], S.V[j * S.ldv + c2]);
          const float on = c0 + u < n ? 1.0f : 0.0f;
          w[u].x += on * t.x;
          w[u].y += on * t.y;
        }
      }
      S.A0[e] = make_float2((w[0].x + w[1].x) + (w[2].x + w[3].x), (w[0].y + w[1].y) + (w[2].y + w[3].y));
    }
    __syncthreads();
    for (int j = tid; j < n; j += nthr) {
      if (j < m) {
        float lam[4] = {0.f, 0.f, 0.f, 0.f};
        for (int i0 = 0; i0 < n; i0 += 4) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int i = min(i0 + u, n - 1);
            const float2 q = S.V[j * S.ldv + i], w = S.A0[j * n + i];
            lam[u] += i0 + u < n ? q.x * w.x + q.y * w.y : 0.0f;
          }
        }
        evu[j] = (lam[0] + lam[1]) + (lam[2] + lam[3]);
      } else {
        const int bcol = rk[j - m];
        evu[j] = Bm[bcol * KW + bcol].x;
      }
    }
    __syncthreads();
  } else {
    // complement eigenvalues as their mean, and its largest one by power steps
    // (wave 0, lane = row): v <- (I - U U^H) A v from pseudo-random phases
    if (tid < 64) {
      const int i = tid;
      float tr = 0.0f, sth = 0.0f;
      for (int c2 = 0; c2 < n; ++c2) tr += A[c2 * ld + c2].x;
      for (int k = 0; k < KW; ++k) sth += Bm[k * KW + k].x;
      const float mu = m > 0 ? (tr - sth) / (float)m : 0.0f;
      float sn, cs;
      __sincosf(6.2831853f * (float)((i * 40503u) & 0xffffu) * (1.0f / 65536.0f), &sn, &cs);
      float2 v = i < n ? make_float2(cs, sn) : make_float2(0.f, 0.f);
      float2 *vt = Y;  // v broadcast through LDS (Y is free without a complement)
      float lam = 0.0f;
      for (int it = 0; it < 4; ++it) {
        // project out the KW Ritz vectors (columns m .. n-1 of V)
        float cr[kMaxWarm], ci[kMaxWarm];
#pragma unroll
        for (int k = 0; k < kMaxWarm; ++k) {
          const float2 u = (k < KW && i < n) ? S.V[(m + k) * S.ldv + i] : make_float2(0.f, 0.f);
          cr[k] = wave_sum(u.x * v.x + u.y * v.y);
          ci[k] = wave_sum(u.x * v.y - u.y * v.x);
        }
#pragma unroll
        for (int k = 0; k < kMaxWarm; ++k) {
          const float2 u = (k < KW && i < n) ? S.V[(m + k) * S.ldv + i] : make_float2(0.f, 0.f);
          v.x -= cr[k] * u.x - ci[k] * u.y;
          v.y -= cr[k] * u.y + ci[k] * u.x;
        }
        if (it == 3) break;
        const float nv = wave_sum(v.x * v.x + v.y * v.y);
        const float inv = nv > 1e-30f ? rsqrtf(nv) : 0.0f;
        v.x *= inv;
        v.y *= inv;
        if (i < n) vt[i] = v;
        wave_sync();
        // w = A v, four partial sums
        float2 w4[4] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
        for (int j0 = 0; j0 < n; j0 += 4) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int j = min(j0 + u, n - 1);
            const float2 t = cmul(A[j * ld + min(i, n - 1)], vt[j]);
            const float on = j0 + u < n ? 1.0f : 0.0f;
            w4[u].x += on * t.x;
            w4[u].y += on * t.y;
          }
        }
        wave_sync();
        float2 w = i < n ? make_float2((w4[0].x + w4[1].x) + (w4[2].x + w4[3].x), (w4[0].y + w4[1].y) + (w4[2].y + w4[3].y))
                         : make_float2(0.f, 0.f);
        lam = wave_sum(v.x * w.x + v.y * w.y);  // Rayleigh quotient of the unit v
        v = w;
      }
      if (tid == 0) S.red[2] = lam;
      for (int j = tid; j < n; j += 64) {
        if (j < m) {
          evu[j] = mu;
        } else {
          const int bcol = rk[j - m];
          evu[j] = Bm[bcol * KW + bcol].x;
        }
      }
    }
    __syncthreads();
    cmax = S.red[2];
  }
  if (tid < n) {
    const float lj = evu[tid];
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += (evu[k] < lj) || (evu[k] == lj && k < tid);
    S.perm[rank] = tid;
    S.ev[rank] = lj;
  }
  __syncthreads();
  // leakage test on the K largest
  if (tid < 64) {
    float leak = 0.0f, rest = 0.0f;
    bool ok = K <= KW;
    for (int k = 0; k < n - K; ++k) rest += S.ev[k];
    rest = (n > K) ? rest / (float)(n - K) : 0.0f;
    // without the complement basis: its largest eigenvalue must stay near the floor
    if (!full && cmax > 1.5f * rest) ok = false;
    for (int k = 0; k < K && ok; ++k) {
      const int col = S.perm[n - 1 - k];
      const float th = S.ev[n - 1 - k];
      if (col < m) {
        // a complement direction among the K largest: harmless at the noise
        // floor, a missed interferer above it
        if (th > 1.5f * rest) ok = false;
        continue;
      }
      const float r2 = wave_sum(tid < n ? S.res[(col - m) * 64 + tid] : 0.0f);
      // interference left by mode k ~ angle^2 x its excess over the noise floor
      leak += th > 0.0f ? r2 / th * fmaxf(th - rest, 0.0f) / th : 1e30f;
    }
    if (tid == 0) {
      S.red[1] = ok ? leak / fmaxf(rest, 1e-30f) : 1e30f;
      S.kdet[1] = (ok && leak <= tol * fmaxf(rest, 1e-30f)) ? 1 : 0;
    }
  }
  __syncthreads();
  return S.kdet[1] != 0;
}

// ---- compensated FP32 ("double-float") arithmetic: value = hi + lo ----
struct DF {
  float hi, lo;
};
__device__ __forceinline__ DF df_two_sum(float a, float b) {
  const float s = a + b, bb = s - a;
  return DF{s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DF df_add(DF a, DF b) {
  const DF s = df_two_sum(a.hi, b.hi);
  const float e = s.lo + (a.lo + b.lo);
  const float hi = s.hi + e;
  return DF{hi, e - (hi - s.hi)};
}
// a * x for a double-float a and a float x: the hi product exact through FMA
__device__ __forceinline__ DF df_mul(DF a, float x) {
  const float p = a.hi * x;
  return DF{p, fmaf(a.hi, x, -p) + a.lo * x};
}
__device__ __forceinline__ DF df_neg(DF a) { return DF{-a.hi, -a.lo}; }
// an integer below 2^48 in magnitude, exactly
__device__ __forceinline__ DF df_from_i64(long long v) {
  const float hi = (float)v;
  return DF{hi, (float)(v - (long long)hi)};
}

// Eigenvalues of the K removed modes as Rayleigh quotients of their computed
// (whitened) eigenvectors u: lambda = u^H S A S u / u^H u with A the block's
// exact integer covariance and S = diag(1 / d) the whitening, in compensated
// FP32.  The FP32 Jacobi (or Rayleigh-Ritz) eigenvalue of a weak mode carries
// the absolute error eps |A_w| of the FP32 whitened matrix, large relative to
// the mode itself when a strong interferer sets |A_w|; the quotient's error is
// second order in the eigenvector error (|r|^2 / gap) and the products of the
// matrix-vector product are exact, so its sums keep the small mode.  The
// diagonal scaling in FP32 perturbs the quotient only relatively (O(eps)).
// Work item (row i, chunk of kJc columns): the chunk's exact entries are
// loaded once for all modes and its terms Re(conj(x_i) A_ij x_j) (x = S u)
// summed per mode; a block reduction in compensated arithmetic finishes.
// Runs between the eigen solve and the outputs (S.A, whose diagonal the sort
// has already copied, is the reduction scratch); overwrites S.ev of the top
// K, up to kRefineModes modes per pass.  cov(i, j) (i <= j) gives the exact
// A(i, j) as a longlong2.
constexpr int kRefineModes = 4;
template <typename Cov>
__device__ __forceinline__ void refine_removed_eigenvalues(EigSmem &S, int n, int K, int whiten, Cov cov) {
  constexpr int kJc = 6;
  const int tid = threadIdx.x, nthr = blockDim.x, lda = jac_ne(n), lane = tid & 63, w = tid >> 6, nw = nthr >> 6;
  const int nch = (n + kJc - 1) / kJc;
  float2 *red = S.A;  // [kRefineModes][waves] (hi, lo)
  for (int m0 = 0; m0 < K; m0 += kRefineModes) {
    const int km = min(kRefineModes, K - m0);
    DF acc[kRefineModes];
#pragma unroll
    for (int m = 0; m < kRefineModes; ++m) acc[m] = DF{0.f, 0.f};
    for (int item = tid; item < n * nch; item += nthr) {
      const int i = item / nch, j0 = (item - i * nch) * kJc;
      DF ar[kJc], ai[kJc];
#pragma unroll
      for (int jj = 0; jj < kJc; ++jj) {
        const int j = min(j0 + jj, n - 1);
        const longlong2 a = cov(min(i, j), max(i, j));
        ar[jj] = df_from_i64(a.x);
        ai[jj] = df_from_i64(i > j ? -a.y : a.y);  // A(i, j) = conj(A(j, i)) below the diagonal
      }
      const float si = whiten ? 1.0f / S.d[i] : 1.0f;
#pragma unroll
      for (int m = 0; m < kRefineModes; ++m) {
        if (m >= km) continue;
        const float2 *u = S.V + S.perm[n - 1 - (m0 + m)] * S.ldv;
        DF yr{0.f, 0.f}, yi{0.f, 0.f};
#pragma unroll
        for (int jj = 0; jj < kJc; ++jj) {
          const int j = j0 + jj;
          if (j >= n) continue;
          const float sj = whiten ? 1.0f / S.d[j] : 1.0f;
          const float xr = u[j].x * sj, xi = u[j].y * sj;
          // (ar + i ai)(xr + i xi)
          yr = df_add(yr, df_mul(ar[jj], xr));
          yr = df_add(yr, df_neg(df_mul(ai[jj], xi)));
          yi = df_add(yi, df_mul(ar[jj], xi));
          yi = df_add(yi, df_mul(ai[jj], xr));
        }
        // Re(conj(x_i) y_i)
        acc[m] = df_add(acc[m], df_add(df_mul(yr, u[i].x * si), df_mul(yi, u[i].y * si)));
      }
    }
    // block reduction per mode: across the wave by DPP moves (the pattern of
    // wave_sum, compensated adds; the total in lane 63), then across waves
#pragma unroll
    for (int m = 0; m < kRefineModes; ++m) {
      if (m < km) {
        DF v = acc[m];
        v = df_add(v, DF{dpp_move<0x111, 0xf>(v.hi), dpp_move<0x111, 0xf>(v.lo)});
        v = df_add(v, DF{dpp_move<0x112, 0xf>(v.hi), dpp_move<0x112, 0xf>(v.lo)});
        v = df_add(v, DF{dpp_move<0x114, 0xf>(v.hi), dpp_move<0x114, 0xf>(v.lo)});
        v = df_add(v, DF{dpp_move<0x118, 0xf>(v.hi), dpp_move<0x118, 0xf>(v.lo)});
        v = df_add(v, DF{dpp_move<0x142, 0xa>(v.hi), dpp_move<0x142, 0xa>(v.lo)});
        v = df_add(v, DF{dpp_move<0x143, 0xc>(v.hi), dpp_move<0x143, 0xc>(v.lo)});
        acc[m] = DF{rlf(v.hi, 63), rlf(v.lo, 63)};
      }
    }
    __syncthreads();  // the previous users of S.A are done
    if (lane == 0) {
#pragma unroll
      for (int m = 0; m < kRefineModes; ++m)
        if (m < km) red[m * nw + w] = make_float2(acc[m].hi, acc[m].lo);
    }
    __syncthreads();
    if (tid < km) {
      DF t{0.f, 0.f};
      for (int ww = 0; ww < nw; ++ww) t = df_add(t, DF{red[tid * nw + ww].x, red[tid * nw + ww].y});
      const float2 *u = S.V + S.perm[n - 1 - (m0 + tid)] * S.ldv;
      float nrm = 0.0f;
      for (int i = 0; i < n; ++i) nrm += u[i].x * u[i].x + u[i].y * u[i].y;
      S.ev[n - 1 - (m0 + tid)] = (t.hi + t.lo) / nrm;
    }
    __syncthreads();
  }
}

__device__ __forceinline__ TriLds tri_lds(float2 *W, int n) {
  TriLds t;
  t.part = W;
  t.fac = reinterpret_cast<float *>(W);
  t.z = t.fac + 4 * kTriK * n;
  float *tab = reinterpret_cast<float *>(W + tri_w_base(n));
  t.ta = tab;
  t.tb = tab + 64;
  t.hc = tab + 128;
  t.ph = reinterpret_cast<float2 *>(tab + 192);
  t.flag = reinterpret_cast<int *>(tab + 320);
  return t;
}

// Inclusive prefix maximum over the lanes (any sign): the DPP pattern of
// wave_sum with -3e38 where a move has no source.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move_neg(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, -3e38f),
                                                               __builtin_bit_cast(int, v), CTRL, ROW_MASK, 0xf,
                                                               false));
}
__device__ __forceinline__ float wave_scan_max(float v) {
  v = fmaxf(v, dpp_move_neg<0x111, 0xf>(v));  // row_shr:1
  v = fmaxf(v, dpp_move_neg<0x112, 0xf>(v));  // row_shr:2
  v = fmaxf(v, dpp_move_neg<0x114, 0xf>(v));  // row_shr:4
  v = fmaxf(v, dpp_move_neg<0x118, 0xf>(v));  // row_shr:8
  v = fmaxf(v, dpp_move_neg<0x142, 0xa>(v));  // row_bcast:15 into rows 1 and 3
  v = fmaxf(v, dpp_move_neg<0x143, 0xc>(v));  // row_bcast:31 into rows 2 and 3
  return v;
}
// Maximum over the 64 lanes (any sign).
__device__ __forceinline__ float wave_max_any(float v) { return rlf(wave_scan_max(v), 63); }

// Waves that own columns of the matrix during the reduction (with 4 waves,
// wave 0 keeps to the serial part).
template <int NW>
constexpr int tri_owners() { return NW >= 4 ? NW - 1 : NW; }

// Reduction of the Hermitian A (n >= 3, full storage, leading dimension
// jac_ne(n)) to T = Q^H A Q tridiagonal with Q = H_0 ... H_{n-3},
// H_j = I - c_j u_j u_j^H acting on rows j+1.. (Hermitian reflectors): H_j
// maps column j below the diagonal, x, to -e^{i arg x_{j+1}} |x| e_{j+1}.
// Step j: y = c A22 u, w = y - (c/2)(u^H y) u, A22 <- A22 - u w^H - w u^H.
// The matrix stays in LDS one update behind: the product of step j uses the
// columns before step j-1's update and is corrected by -u'(w'^H u) -
// w'(u'^H u) (u', w' of step j-1); the owner waves apply that update to the
// columns above j+1 while wave 0 forms y, w, the next column (column j+1 with
// both pending updates) and the next reflector.  Owners walk their columns
// in loops whose loads batch, the vectors' entries read as LDS broadcasts
// (no cross-lane moves).  Two barriers per step.  u_j is left in column j
// of A (rows j+1..), the tridiagonal in T: ta (real diagonal), tb |e_j|, ph
// the phase of e_j, hc c_j.
template <int NTHR>
__device__ __forceinline__ void tri_reduce(float2 *A, int n, const TriLds &T) {
  constexpr int NW = NTHR / 64, OW = tri_owners<NW>();
  const int i = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lda = jac_ne(n);
  const int ow = NW >= 4 ? w - 1 : w;    // owner index (-1: wave 0 with 4 waves)
  float2 *pp = T.part;                   // [OW][64] partial products
  float4 *uw = reinterpret_cast<float4 *>(pp + OW * 64);  // [2][64] (u_j, w_j) by parity
  float2 *un = reinterpret_cast<float2 *>(uw + 2 * 64);   // [2][64] u_j by parity (the product's)
  float2 *sink = un + 2 * 64;            // [64] stores of the lanes past n
  // wave 0's vectors: u_j, u_{j-1}, w_{j-1} (lane i: row i)
  float2 uc = make_float2(0.f, 0.f), up = make_float2(0.f, 0.f), wp = make_float2(0.f, 0.f);
  // wave 0: the reflector of column x (rows > j) for step j, into un / A / T
  auto reflector = [&](int j, float2 x, float dj) {
    const float s = wave_sum(i > j + 1 ? x.x * x.x + x.y * x.y : 0.0f);
    const float2 al = rl2(x, j + 1);
    const float a2 = al.x * al.x + al.y * al.y;
    const float av = __builtin_amdgcn_sqrtf(a2), nx = __builtin_amdgcn_sqrtf(a2 + s);
    const bool refl = s > 0.0f;
    const float ia = av > 0.0f ? __builtin_amdgcn_rcpf(av) : 0.0f;
    const float2 ph = av > 0.0f ? make_float2(al.x * ia, al.y * ia) : make_float2(1.f, 0.f);
    const float c = refl ? __builtin_amdgcn_rcpf(nx * (nx + av)) : 0.0f;
    float2 u = i == j + 1 ? make_float2(ph.x * (av + nx), ph.y * (av + nx)) : x;
    if (i <= j || i >= n) u = make_float2(0.f, 0.f);
    un[(j & 1) * 64 + i] = u;
    if (i > j && i < n) A[j * lda + i] = u;
    if (i == 0) {
      T.ta[j] = dj;
      T.tb[j] = refl ? nx : av;
      T.ph[j] = refl ? make_float2(-ph.x, -ph.y) : ph;
      T.hc[j] = c;
    }
    return u;
  };
  if (w == 0) uc = reflector(0, i > 0 && i < n ? A[i] : make_float2(0.f, 0.f), A[0].x);
  __syncthreads();
  // this owner's first column >= lo
  auto first_col = [&](int lo) {
    const int m = NW >= 4 ? ow : w;  // columns c = m (mod OW)
    return lo + ((m - lo) % OW + OW) % OW;
  };
  for (int j = 0; j + 2 < n; ++j) {
    // wave 0, while the owners form the product: the sums and the part of the
    // next column that do not need the product
    float al_r = 0.f, al_i = 0.f, be_r = 0.f, be_i = 0.f, dn0 = 0.f;
    float2 u1 = make_float2(0.f, 0.f), xq = make_float2(0.f, 0.f);
    if (w == 0) {
      al_r = wave_sum(wp.x * uc.x + wp.y * uc.y);
      al_i = wave_sum(wp.x * uc.y - wp.y * uc.x);
      be_r = wave_sum(up.x * uc.x + up.y * uc.y);
      be_i = wave_sum(up.x * uc.y - up.y * uc.x);
      // column j+1 (before step j-1's update) with step j-1's update
      const float2 xo = i < n ? A[(j + 1) * lda + i] : make_float2(0.f, 0.f);
      const float2 u0 = rl2(up, j + 1), w0 = rl2(wp, j + 1);
      u1 = rl2(uc, j + 1);
      const float2 t3 = cmulc(up, w0), t4 = cmulc(wp, u0);
      xq = make_float2(xo.x - t3.x - t4.x, xo.y - t3.y - t4.y);
      dn0 = rlf(xo.x, j + 1) - 2.0f * (u0.x * w0.x + u0.y * w0.y);
    }
    if (ow >= 0) {
      // this wave's part of A22 u_j (columns above j, one update behind)
      // (rows past n read row n-1: finite values that only meet zero entries
      // of u, w and x; no exec-mask branch around the loads)
      const float2 *uj = un + (j & 1) * 64;
      const int ir = min(i, n - 1);
      float2 yp = make_float2(0.f, 0.f);
#pragma unroll 4
      for (int c = first_col(j + 1); c < n; c += OW) {
        const float2 e = uj[c];
        const float2 av = A[c * lda + ir];
        yp.x += av.x * e.x - av.y * e.y;
        yp.y += av.x * e.y + av.y * e.x;
      }
      pp[ow * 64 + i] = yp;
    }
    __syncthreads();
    if (w == 0) {
      const float c = T.hc[j];
      float2 y = make_float2(0.f, 0.f);
#pragma unroll
      for (int q = 0; q < OW; ++q) {
        const float2 t2 = pp[q * 64 + i];
        y.x += t2.x;
        y.y += t2.y;
      }
      // y = c (P - u' al - w' be) with al = w'^H u, be = u'^H u;
      // Re(u^H y) / c = Re(u^H P) - 2 Re(conj(be) al)
      const float g0 = wave_sum(uc.x * y.x + uc.y * y.y);
      y.x -= up.x * al_r - up.y * al_i + wp.x * be_r - wp.y * be_i;
      y.y -= up.x * al_i + up.y * al_r + wp.x * be_i + wp.y * be_r;
      y.x *= c;
      y.y *= c;
      const float kap = 0.5f * c * c * (g0 - 2.0f * (be_r * al_r + be_i * al_i));
      const float2 wv = i > j && i < n ? make_float2(y.x - kap * uc.x, y.y - kap * uc.y) : make_float2(0.f, 0.f);
      uw[(j & 1) * 64 + i] = make_float4(uc.x, uc.y, wv.x, wv.y);
      // the next x and diagonal entry: step j's update on top
      const float2 w1 = rl2(wv, j + 1);
      const float2 t1 = cmulc(uc, w1), t2 = cmulc(wv, u1);
      const float2 xn = i > j + 1 && i < n ? make_float2(xq.x - t1.x - t2.x, xq.y - t1.y - t2.y)
                                           : make_float2(0.f, 0.f);
      const float dn = dn0 - 2.0f * (u1.x * w1.x + u1.y * w1.y);
      up = uc;
      wp = wv;
      if (j + 3 < n) {
        uc = reflector(j + 1, xn, dn);
      } else {  // the last 2 x 2: a_{n-2,n-2} and e_{n-2}
        if (i == 0) {
          T.ta[n - 2] = dn;
          T.hc[n - 2] = 0.0f;
        }
        if (i == n - 1) {
          const float ae = sqrtf(xn.x * xn.x + xn.y * xn.y);
          T.tb[n - 2] = ae;
          T.ph[n - 2] = ae > 0.0f ? make_float2(xn.x / ae, xn.y / ae) : make_float2(1.f, 0.f);
        }
      }
    }
    // step j-1's update on the owners' columns above j+1 (beside wave 0's part)
    if (ow >= 0 && j > 0) {
      const float4 *uwp = uw + ((j - 1) & 1) * 64;
      const float4 me = uwp[i];  // (u'_i, w'_i)
      const int ir = min(i, n - 1);
#pragma unroll 4
      for (int c = first_col(j + 2); c < n; c += OW) {
        const float4 e = uwp[c];
        // a -= u'_i conj(w'_c) + w'_i conj(u'_c) (lanes past n: into the sink)
        const float2 av = A[c * lda + ir];
        const float2 nv = make_float2(av.x - (me.x * e.z + me.y * e.w + me.z * e.x + me.w * e.y),
                                      av.y - (me.y * e.z - me.x * e.w + me.w * e.x - me.z * e.y));
        float2 *dst = i < n ? A + c * lda + i : sink + i;
        *dst = nv;
      }
    }
    __syncthreads();
  }
  // a_{n-1,n-1}: step n-4's update is in, step n-3's (wave 0's u', w') is applied here
  if (w == 0) {
    const float2 un1 = rl2(up, n - 1), wn1 = rl2(wp, n - 1);
    const float d = A[(n - 1) * lda + n - 1].x - 2.0f * (un1.x * wn1.x + un1.y * wn1.y);
    if (i == 0) T.ta[n - 1] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // diagonal similarity D (phi_0 = 1, phi_{j+1} = phi_j e_j / |e_j|): D^H T D
    // is real with off-diagonal |e_j|
    float2 cur = make_float2(1.f, 0.f);
    for (int j = 0; j < n; ++j) {
      const float2 pe = j + 1 < n ? T.ph[j] : make_float2(1.f, 0.f);
      T.ph[j] = cur;
      cur = cmul(cur, pe);
      const float s2 = rsqrtf(cur.x * cur.x + cur.y * cur.y);
      cur.x *= s2;
      cur.y *= s2;
    }
  }
  __syncthreads();
}

// Eigenvalues of the real tridiagonal (ta, tb) into ev, ascending (perm =
// identity): the top nsel by multisection, eigenvalue k in one group of
// lanes of a wave, each lane a Sturm count at its point (the number of
// eigenvalues below it, from the LDL^T pivots with the usual minimum
// pivot); the others (nsel < n) as their mean, trace minus the computed
// ones over n - nsel.  Returns the Gershgorin bound of |T|.
template <int NTHR>
__device__ __forceinline__ float tri_eigenvalues(int n, int nsel, const TriLds &T, float *ev, int *perm) {
  constexpr int NW = NTHR / 64;
  const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
  const float bl = r >= 1 && r < n ? T.tb[r - 1] : 0.f, br = r + 1 < n ? T.tb[r] : 0.f;
  const float ar = r < n ? T.ta[r] : 0.f;
  float glo = -wave_max_any(r < n ? -(ar - bl - br) : -3e38f);
  float ghi = wave_max_any(r < n ? ar + bl + br : -3e38f);
  const float tn = fmaxf(fabsf(glo), fabsf(ghi));
  const float pivmin = 1e-30f * fmaxf(1.0f, wave_max(br * br));
  const float pad = 2.4e-7f * tn * (float)n + pivmin;
  glo -= pad;
  ghi += pad;
  const int k0 = n - nsel;
  const int G = (nsel + NW - 1) / NW, nt = min(63, max(1, 64 / max(G, 1)));
  const int g = r / nt, t = r - g * nt, k = k0 + w * G + g;
  const bool act = nsel > 0 && g < G && k < n;
  int iters = 1;
  for (long p = 1; p < (1L << 30); p *= nt + 1) ++iters;
  const float inv = 1.0f / (float)(nt + 1), fr = (float)(t + 1) * inv;
  float lo = glo, hi = ghi;
  if (nsel > 0 && w * G < nsel) {
    // pivots q_i = (a_i - x) - b_{i-1}^2 / q_{i-1}; with b^2 kept >= 1e-30 a
    // zero pivot gives -inf next (counted) and then a_i - x again, the count
    // of the limit q -> +0, so the chain needs no pivot guard
    (void)pivmin;
    for (int it = 0; it < iters; ++it) {
      const float x = fmaf(hi - lo, fr, lo);
      float q = T.ta[0] - x;
      int cnt = q < 0.0f;
#pragma unroll 4
      for (int i = 1; i < n; ++i) {
        const float b = T.tb[i - 1];
        q = (T.ta[i] - x) - fmaxf(b * b, 1e-30f) * __builtin_amdgcn_rcpf(q);
        cnt += q < 0.0f;
      }
      const unsigned long long bal = __ballot(act && cnt >= k + 1);
      const unsigned long long bits = (bal >> (g * nt)) & ((nt >= 64 ? 0ull : (1ull << nt)) - 1ull);
      const int t1 = bits ? (int)__builtin_ctzll(bits) : nt;
      const float nlo = t1 > 0 ? fmaf(hi - lo, (float)t1 * inv, lo) : lo;
      const float nhi = t1 < nt ? fmaf(hi - lo, (float)(t1 + 1) * inv, lo) : hi;
      lo = nlo;
      hi = nhi;
    }
  }
  if (act && t == 0) ev[k] = 0.5f * (lo + hi);
  if (threadIdx.x < n) perm[threadIdx.x] = threadIdx.x;
  __syncthreads();
  if (w == 0) {
    // ascending through ties (prefix maximum); the uncomputed ones: their mean
    float v = r >= k0 && r < n ? ev[r] : -3e38f;
    const float top = wave_sum(r >= k0 && r < n ? ev[r] : 0.0f), tr = wave_sum(ar);
    const float mean = k0 > 0 ? (tr - top) / (float)k0 : 0.0f;
    v = wave_scan_max(v);
    if (r < n) ev[r] = r < k0 ? fminf(mean, k0 < n ? ev[k0] : mean) : v;
  }
  __syncthreads();
  return tn;
}

// Eigenvectors of the `need` largest eigenvalues into columns n-1, n-2, ...
// of V (the other columns zero), on the real tridiagonal, then
// v = H_0 ... H_{n-3} D z.  An isolated eigenvalue (gap to its neighbours at
// least 1e-3 of the Gershgorin bound) takes the twisted factorization of
// T - lambda I: forward pivots D+ and backward pivots D- (two lanes, two
// short independent chains), the twist r = argmin |D+_r + D-_r - (a_r - lambda)|,
// then z_r = 1, z_i = -(b_i / D+_i) z_{i+1} upwards and
// z_{i+1} = -(b_i / D-_{i+1}) z_i downwards.  The others take inverse
// iteration (two solves of (T - lambda I) z = b, Gaussian elimination with
// partial pivoting, pseudo-random start).  Pivots are kept away from zero;
// modified Gram-Schmidt over the vectors.  Returns false (uniform) if a
// vector is not finite.
template <int NTHR>
__device__ __forceinline__ bool tri_vectors(const float2 *A, int n, int need, const TriLds &T, const float *ev,
                                            float tn, float2 *V, int ldv) {
  constexpr int NW = NTHR / 64;
  const int tid = threadIdx.x, r = tid & 63, w = tid >> 6, lda = jac_ne(n);
  const float tiny = fmaxf(2.4e-7f * tn, 1e-20f);  // a zero matrix: z = 1e20 b
  if (tid == 0) T.flag[0] = 1;
  int *iso = T.flag + 16, *rtw = T.flag + 24;  // [kTriK] isolated, twist index
  if (w == 0) {
    if (r < need) {
      const int q = n - 1 - r;
      const float glo = q >= 1 ? ev[q] - ev[q - 1] : 3e38f, ghi = q + 1 < n ? ev[q + 1] - ev[q] : 3e38f;
      iso[r] = fminf(glo, ghi) >= 1e-3f * tn ? 1 : 0;
    }
    wave_sync();
    // twisted pivots: lane 32 + 2k forward, 33 + 2k backward (D+ and D- of
    // vector k at fac rows 4i and 4i + 1)
    const int kt = (r - 32) >> 1;
    const bool twl = r >= 32 && kt < need && iso[kt & (kTriK - 1)];
    if (twl) {
      const float sig = ev[n - 1 - kt];
      const float *__restrict__ ta = T.ta;
      const float *__restrict__ tb = T.tb;
      float *__restrict__ dp = T.fac + kt;
      auto guard = [&](float d) { return fabsf(d) < tiny ? (d >= 0.0f ? tiny : -tiny) : d; };
      if (((r - 32) & 1) == 0) {
        float d = guard(ta[0] - sig);
        dp[0] = d;
#pragma unroll 4
        for (int i = 0; i + 1 < n; ++i) {
          const float b = tb[i];
          d = guard((ta[i + 1] - sig) - b * b * __builtin_amdgcn_rcpf(d));
          dp[(i + 1) * 4 * kTriK] = d;
        }
      } else {
        float d = guard(ta[n - 1] - sig);
        dp[((n - 1) * 4 + 1) * kTriK] = d;
#pragma unroll 4
        for (int i = n - 2; i >= 0; --i) {
          const float b = tb[i];
          d = guard((ta[i] - sig) - b * b * __builtin_amdgcn_rcpf(d));
          dp[(i * 4 + 1) * kTriK] = d;
        }
      }
    }
    const int k = r;
    const bool own = k < need && !iso[k & (kTriK - 1)];
    float *z = T.z + k, *f = T.fac + k;
    if (own) {
      const float sig = ev[n - 1 - k];
      const unsigned mul = 0x9E3779B1u + 0x632BE5ABu * (unsigned)k;
      // distinct arrays (restrict: the loads of a step need not wait for the
      // previous step's stores)
      const float *__restrict__ ta = T.ta;
      const float *__restrict__ tb = T.tb;
      float *__restrict__ zz = T.z + k;
      float *__restrict__ ff = T.fac + k;
      bool fin = true;
      float sc = 1.0f;  // the right-hand side of the second solve: the first solution over its norm
      for (int it = 0; it < 2; ++it) {
        auto rhs = [&](int ii) {
          return it == 0 ? (float)(((unsigned)(ii + 1) * mul) >> 8) * (1.0f / 8388608.0f) - 1.0f
                         : zz[ii * kTriK] * sc;
        };
        float p = ta[0] - sig, q = tb[0], rr = rhs(0);
#pragma unroll 4
        for (int i = 0; i + 1 < n; ++i) {
          const float bi = tb[i], dn = ta[i + 1] - sig, bn = i + 2 < n ? tb[i + 1] : 0.f;
          const float rn = rhs(i + 1);
          float *fi = ff + i * 4 * kTriK;
          const bool sw = fabsf(bi) > fabsf(p);  // row i+1 is the pivot row
          const float pv = sw ? bi : (fabsf(p) < tiny ? (p >= 0.0f ? tiny : -tiny) : p);
          const float ip = __builtin_amdgcn_rcpf(pv), mlt = (sw ? p : bi) * ip;
          fi[0] = ip;
          fi[kTriK] = sw ? dn : q;
          fi[2 * kTriK] = sw ? bn : 0.0f;
          fi[3 * kTriK] = sw ? rn : rr;
          const float np = sw ? q - mlt * dn : dn - mlt * q;
          const float nq = sw ? -mlt * bn : bn;
          const float nr = sw ? rr - mlt * rn : rn - mlt * rr;
          p = np;
          q = nq;
          rr = nr;
        }
        if (fabsf(p) < tiny) p = p >= 0.0f ? tiny : -tiny;
        float *fl = ff + (n - 1) * 4 * kTriK;
        fl[0] = __builtin_amdgcn_rcpf(p);
        fl[kTriK] = 0.0f;
        fl[2 * kTriK] = 0.0f;
        fl[3 * kTriK] = rr;
        float z1 = 0.f, z2 = 0.f;
        double ss = 0.0;  // |z| up to ~1e20 |b|
#pragma unroll 4
        for (int i = n - 1; i >= 0; --i) {
          const float *fi = ff + i * 4 * kTriK;
          const float zi = (fi[3 * kTriK] - fi[kTriK] * z1 - fi[2 * kTriK] * z2) * fi[0];
          zz[i * kTriK] = zi;
          ss += (double)zi * zi;
          z2 = z1;
          z1 = zi;
        }
        fin = fin && ss > 0.0 && ss < 1e300;
        sc = fin ? (float)(1.0 / sqrt(ss)) : 0.0f;
      }
#pragma unroll 4
      for (int i = 0; i < n; ++i) zz[i * kTriK] *= sc;
      if (!fin) T.flag[0] = 0;
    }
    wave_sync();
    // the twist of each isolated vector
    for (int kk = 0; kk < need; ++kk) {
      if (iso[kk]) {
        const float sig = ev[n - 1 - kk];
        const float g = r < n ? fabsf(T.fac[(r * 4) * kTriK + kk] + T.fac[(r * 4 + 1) * kTriK + kk] - (T.ta[r] - sig))
                              : 3e38f;
        const float gm = -wave_max_any(-g);
        const unsigned long long bal = __ballot(r < n && g == gm);
        if (r == 0) rtw[kk] = bal ? (int)__builtin_ctzll(bal) : 0;
      }
    }
    wave_sync();
    if (twl) {
      const int rt = rtw[kt];
      const float *__restrict__ tb = T.tb;
      const float *__restrict__ dp = T.fac + kt;
      float *__restrict__ zz = T.z + kt;
      float zc = 1.0f;
      if (((r - 32) & 1) == 0) {
        zz[rt * kTriK] = 1.0f;
        for (int i = rt - 1; i >= 0; --i) {
          zc = -(tb[i] * __builtin_amdgcn_rcpf(dp[(i * 4) * kTriK])) * zc;
          zz[i * kTriK] = zc;
        }
      } else {
        for (int i = rt; i + 1 < n; ++i) {
          zc = -(tb[i] * __builtin_amdgcn_rcpf(dp[((i + 1) * 4 + 1) * kTriK])) * zc;
          zz[(i + 1) * kTriK] = zc;
        }
      }
    }
    wave_sync();
    for (int kk = 0; kk < need; ++kk) {
      if (iso[kk]) {
        const float v = r < n ? T.z[r * kTriK + kk] : 0.0f;
        const float ss = wave_sum(v * v);
        const bool fin = ss > 0.0f && ss < 3e38f;
        if (r < n) T.z[r * kTriK + kk] = fin ? v * rsqrtf(ss) : 0.0f;
        if (r == 0 && !fin) T.flag[0] = 0;
      }
    }
    // modified Gram-Schmidt, descending eigenvalues (vector k after 0 .. k-1)
    for (int j = 0; j + 1 < need; ++j) {
      wave_sync();
      if (own && k > j) {
        const float *zj = T.z + j;
        float dot = 0.0f;
        for (int i = 0; i < n; ++i) dot += zj[i * kTriK] * z[i * kTriK];
        float ss = 0.0f;
        for (int i = 0; i < n; ++i) {
          const float v = z[i * kTriK] - dot * zj[i * kTriK];
          z[i * kTriK] = v;
          ss += v * v;
        }
        const float sc = ss > 0.0f ? rsqrtf(ss) : 0.0f;
        if (k == j + 1)
          for (int i = 0; i < n; ++i) z[i * kTriK] *= sc;
      }
    }
  }
  for (int e = tid; e < (n - need) * n; e += NTHR) {
    const int col = e / n, i = e - col * n;
    V[col * ldv + i] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  for (int k = w; k < need; k += NW) {
    const float zi = r < n ? T.z[r * kTriK + k] : 0.f;
    const float2 phr = r < n ? T.ph[r] : make_float2(0.f, 0.f);
    float2 v = make_float2(phr.x * zi, phr.y * zi);
    for (int j = n - 3; j >= 0; --j) {
      const float c = T.hc[j];  // 0: no reflector (the product vanishes)
      const float2 uj = r > j && r < n ? A[j * lda + r] : make_float2(0.f, 0.f);
      const float dr = wave_sum(uj.x * v.x + uj.y * v.y), di = wave_sum(uj.x * v.y - uj.y * v.x);
      const float2 s = make_float2(c * dr, c * di);  // c u^H v
      v.x -= s.x * uj.x - s.y * uj.y;
      v.y -= s.x * uj.y + s.y * uj.x;
    }
    if (r < n) V[(n - 1 - k) * ldv + r] = v;
  }
  __syncthreads();
  const bool ok = T.flag[0] != 0;
  __syncthreads();
  return ok;
}

// MFMA A-operand fragments of beam rows [row0, row0 + nrows) of one (channel,
// pol) in the layout [NMT][NK][hi, lo][64 lanes] x 8 halves that the beam
// kernel reads (rows of [Re; Im] pairs, K = (antenna, re/im)); weight(e, a)
// gives the complex weight of beam e (0-based within the rows) on antenna a.
// Each lane of a fragment holds one row, so writers of different rows never
// share a fragment lane; rows no writer owns stay as allocated (zero).
template <typename Weight>
__device__ __forceinline__ void write_beam_operands(v8h *wop_cp, int NMT, int NK, int wprec, int row0, int nrows,
                                                    Weight weight) {
  for (int idx = threadIdx.x; idx < NMT * NK * 64; idx += blockDim.x) {
    const int lane = idx & 63, j = (idx >> 6) % NK, mt = (idx >> 6) / NK;
    const int m = 32 * mt + (lane & 31), kg = lane >> 5;
    const int e = (m >> 1) - row0, u = m & 1;
    if (e < 0 || e >= nrows) continue;
    v8h hi, lo;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float2 w = weight(e, 8 * j + 4 * kg + i);
      const float v0 = u == 0 ? w.x : w.y;
      const float v1 = u == 0 ? -w.y : w.x;
      _Float16 h0, l0, h1, l1;
      weight_operands(v0, wprec, h0, l0);
      weight_operands(v1, wprec, h1, l1);
      hi[2 * i] = h0;
      lo[2 * i] = l0;
      hi[2 * i + 1] = h1;
      lo[2 * i + 1] = l1;
    }
    v8h *dst = wop_cp + ((long)mt * NK + j) * 2 * 64;
    dst[lane] = hi;
    dst[64 + lane] = lo;
  }
}

// One (channel, pol) eigen / RFI-null problem in this workgroup.
// WARM = false: plain Jacobi; WARM = true: warm-start solve, a rejected
// problem re-solved with the plain solver in the same workgroup.
template <int NTHR, bool WARM>
__device__ __forceinline__ void eig_null_problem(const EigParams &P, char *smem_raw, int c, int p) {
  const int n = P.R, B = P.B;
  EigSmem S = carve(smem_raw, n, B, WARM, NTHR / 64);
  const int tid = threadIdx.x, nthr = blockDim.x, lda = jac_ne(n);
  const long nbl = (long)n * (n + 1) / 2;

  // 1. dense Hermitian covariance of pol p (A(i,j) = V[bl(i,j)][p][p], i <= j)
  // 2. whitening (spatial.cuh:1065-1093); off: the raw covariance, as the
  //    pulsar-fold pipeline's RFI branch decomposes it (d = 1 makes the
  //    unwhitening below a no-op)
  // the beam weights, loaded now so their latency hides behind the solve
  constexpr int kWPre = 3;
  const bool wfit = B * n <= kWPre * nthr;
  __half2 wpre[kWPre];
#pragma unroll
  for (int q = 0; q < kWPre; ++q) {
    const int idx = tid + q * nthr;
    wpre[q] = wfit && idx < B * n ? reinterpret_cast<const __half2 *>(P.weights)[((long)(c * 2 + p) * B) * n + idx]
                                  : __half2{};
  }
  auto load_matrix = [&]() {
    for (int idx = tid; idx < nbl; idx += nthr) {
      int j = (int)((sqrtf(8.0f * idx + 1.0f) - 1.0f) * 0.5f);
      while ((long)j * (j + 1) / 2 > idx) --j;
      while ((long)(j + 1) * (j + 2) / 2 <= idx) ++j;
      const int i = idx - j * (j + 1) / 2;
      long re = 0, im = 0;
      for (int k = 0; k < P.nsplit; ++k) {
        const int2 v = reinterpret_cast<const int2 *>(P.vis_part)[(((long)k * P.C + c) * 4 + 3 * p) * nbl + idx];
        re += v.x;
        im += v.y;
      }
      const float2 a = make_float2((float)re, (float)im);
      S.A[j * lda + i] = a;
      S.A[i * lda + j] = make_float2(a.x, -a.y);
    }
    __syncthreads();
    if (tid < n) S.d[tid] = P.whiten ? sqrtf(fmaxf(S.A[tid * lda + tid].x, 1e-10f)) : 1.0f;
    __syncthreads();
    if (P.whiten) {
      for (int i = tid; i < n * n; i += nthr) {
        const int col = i / n, row = i - col * n;
        const float den = S.d[row] * S.d[col];
        S.A[col * lda + row].x /= den;
        S.A[col * lda + row].y /= den;
      }
      __syncthreads();
    }
  };
  load_matrix();
  // 3. eigen decomposition: plain cyclic Jacobi, or a warm start refined by a
  //    few block sweeps and accepted only if the top-K residual is within tolerance
  const long cp = (long)c * 2 + p;
  const int fixed = P.fixed_k[c * 2 + p];
  int sweeps = 0, K = 0, fell = 0;
  if constexpr (WARM) {
    const int KW = P.KW;
    for (int e = tid; e < KW * n; e += nthr) S.Y[e] = P.proposal[cp * P.prop_stride + e];
    __syncthreads();
    K = max(0, min(n, fixed));
    const bool ok = warm_solve(S.A, lda, n, KW, K, P.refine_sweeps, P.res_tol, S, P.full_outputs || P.shrink);
    sweeps = P.refine_sweeps;
    if (ok) {
      if (tid == 0) S.kdet[0] = K;
      __syncthreads();
    } else {
      // rejected: the plain solver on the same matrix (the warm solve leaves A
      // untouched), in this workgroup, so no second launch is needed
      sweeps += jacobi_lds<NTHR>(S.A, S.V, n, P.max_sweeps, P.tol, S.jl());
      sort_eigen(S.A, n, S.perm, S.ev);
      K = select_k(S.ev, n, fixed, P.detect, P.delta, S.kdet);
      fell = 1;
    }
  } else {
    // without eigen outputs only the removed modes' vectors are needed: the
    // tridiagonal path, unless it needs more than kTriK vectors (or one is
    // not finite): then the matrix is formed again for the Jacobi solver
    bool solved = false;
    if (!P.full_outputs && !P.shrink && n >= 3) {
      const TriLds T = tri_lds(S.W, n);
      tri_reduce<NTHR>(S.A, n, T);
      // with detection all eigenvalues (percentiles); else the top ones used
      const int kw = P.basis_out ? min(P.KW, n) : 0;
      // (one more than the vectors need: the gap below the last one)
      const int nsel = P.detect ? n : min(n, max(max(0, min(n, fixed)), kw) + 1);
      const float tn = tri_eigenvalues<NTHR>(n, nsel, T, S.ev, S.perm);
      K = select_k(S.ev, n, fixed, P.detect, P.delta, S.kdet);
      const int need = max(K, kw);
      if (need <= kTriK) solved = tri_vectors<NTHR>(S.A, n, need, T, S.ev, tn, S.V, S.ldv);
      if (!solved) load_matrix();
    }
    if (!solved) {
      sweeps = jacobi_lds<NTHR>(S.A, S.V, n, P.max_sweeps, P.tol, S.jl());
      sort_eigen(S.A, n, S.perm, S.ev);
      K = select_k(S.ev, n, fixed, P.detect, P.delta, S.kdet);
    }
  }
  // 4. eigenvalues of the removed modes to the precision of the exact covariance
  if (K > 0)
    refine_removed_eigenvalues(S, n, K, P.whiten, [&](int i, int j) {
      const long bl = (long)j * (j + 1) / 2 + i;
      long long re = 0, im = 0;
      for (int k = 0; k < P.nsplit; ++k) {
        const int2 v = reinterpret_cast<const int2 *>(P.vis_part)[(((long)k * P.C + c) * 4 + 3 * p) * nbl + bl];
        re += v.x;
        im += v.y;
      }
      return make_longlong2(re, im);
    });
  // modes clearly above the noise floor (> 4x the mean of the remaining eigenvalues):
  // the label count of the warm-start training set
  if (P.nsig && tid < 64) {
    const float rest = wave_sum(tid < n - K ? S.ev[tid] : 0.0f);
    const float mean = n > K ? rest / (float)(n - K) : 0.0f;
    const float ns = wave_sum(tid >= n - K && tid < n && S.ev[tid] > 4.0f * mean ? 1.0f : 0.0f);
    if (tid == 0) P.nsig[cp] = (int)ns;
  }
  // top-KW whitened eigenvectors (descending): the next block's warm start / training labels
  if (P.basis_out) {
    for (int e = tid; e < P.KW * n; e += nthr) {
      const int k = e / n, i = e - k * n;
      P.basis_out[cp * P.KW * n + e] = k < n ? S.V[S.perm[n - 1 - k] * S.ldv + i] : make_float2(0.f, 0.f);
    }
  }
  if (tid == 0 && P.fallback) P.fallback[cp] = fell;
  // FP64 orthogonality check of the null basis: max |u_a^H u_b - delta_ab| over the
  // K removed (whitened) eigenvectors, one (a, b) pair per wave (lanes: rows)
  if (P.orth) {
    double e = 0.0;
    const int lane = tid & 63, wq = tid >> 6, nwq = nthr >> 6;
    for (int ab = wq; ab < K * K; ab += nwq) {
      const int a = ab / K, b = ab - a * K;
      const int ca = S.perm[n - 1 - a], cb = S.perm[n - 1 - b];
      const float2 x = lane < n ? S.V[ca * S.ldv + lane] : make_float2(0.f, 0.f);
      const float2 y = lane < n ? S.V[cb * S.ldv + lane] : make_float2(0.f, 0.f);
      double re = wave_sum_f64((double)x.x * y.x + (double)x.y * y.y);
      const double im = wave_sum_f64((double)x.x * y.y - (double)x.y * y.x);
      if (a == b) re -= 1.0;
      e = fmax(e, sqrt(re * re + im * im));
    }
    const float ef = block_max((float)e, S.red);
    if (tid == 0) P.orth[cp] = ef;
  }
  // 5. output basis U (sorted columns) into S.A: unwhitened+normalised (null) or whitened (shrink);
  //    without eigen outputs the null needs only the K removed columns.  One wave per column.
  {
    const int lane = tid & 63, wq = tid >> 6, nwq = nthr >> 6;
    const int kfirst = (P.full_outputs || P.shrink) ? 0 : n - K;
    for (int col = kfirst + wq; col < n; col += nwq) {
      const int src = S.perm[col];
      float2 v = lane < n ? S.V[src * S.ldv + lane] : make_float2(0.f, 0.f);
      if (!P.shrink) {
        const float dl = lane < n ? S.d[lane] : 0.0f;
        v.x *= dl;
        v.y *= dl;
        const float nrm = wave_sum(v.x * v.x + v.y * v.y);
        const float inv = nrm > 0.0f ? rsqrtf(nrm) : 0.0f;
        v.x *= inv;
        v.y *= inv;
      }
      if (lane < n) S.A[col * n + lane] = v;
    }
  }
  if (tid < n) {
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
  // weights of this (c, p): rows 0..B-1 original (prefetched at the start)
  if (wfit) {
#pragma unroll
    for (int q = 0; q < kWPre; ++q)
      if (tid + q * nthr < B * n) S.W[tid + q * nthr] = __half22float2(wpre[q]);
  } else {
    for (int idx = tid; idx < B * n; idx += nthr) {
      const __half2 w = reinterpret_cast<const __half2 *>(P.weights)[((long)(c * 2 + p) * B) * n + idx];
      S.W[idx] = __half22float2(w);
    }
  }
  __syncthreads();
  // 6. eigen outputs (when requested)
  if (P.full_outputs) {
    if (tid < n) P.evals[cp * n + tid] = S.ev[tid];
    for (int i = tid; i < n * n; i += nthr)
      reinterpret_cast<float2 *>(P.evecs)[cp * n * n + i] = S.A[i];
  }
  if (tid == 0) {
    P.counts[cp] = K;
    if (P.sweeps) P.sweeps[cp] = sweeps;
  }
  // 7. W' = W P with P = I - U_K U_K^H (null) or V diag(f) V^H (shrink):
  //    G(b,k) = sum_i W(b,i) u_k(i);  W'(b,i) = [null] W(b,i) + sum_k f_k G(b,k) conj(u_k(i))
  //    null: f_k = 0 below the K removed modes, so only k >= n - K take part
  //    (same sums, same order: the skipped terms are exact zeros)
  const int klo = P.shrink ? 0 : n - K, nk = n - klo;
  for (int idx = tid; idx < B * nk; idx += nthr) {
    const int b = idx / nk, k = klo + idx % nk;
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
    for (int k = klo; k < n; ++k) {
      const float2 t = cmulc(S.G[b * n + k], S.A[k * n + i]);
      acc.x += t.x;
      acc.y += t.y;
    }
    S.W[(B + b) * n + i] = acc;
  }
  __syncthreads();
  // 8. MFMA operands of the 2B beam rows [W; W'] (split hi + lo FP16, FP16 or BF16)
  write_beam_operands(P.wop + cp * P.NMT * P.NK * 2 * 64, P.NMT, P.NK, P.wprec, 0, 2 * B,
                      [&](int e, int a) { return a < n ? S.W[e * n + a] : make_float2(0.f, 0.f); });
}

template <int NTHR, bool WARM>
__global__ void __launch_bounds__(NTHR) eig_null_kernel(EigParams P) {
  extern __shared__ __align__(16) char smem_raw[];
  eig_null_problem<NTHR, WARM>(P, smem_raw, blockIdx.x >> 1, blockIdx.x & 1);
}

// The same solve for blocks that take the tridiagonal path (no eigen outputs, no
// shrink, n >= 3; its rare Jacobi fallback still runs here), held to 96 VGPRs (five
// waves per SIMD): three of these workgroups then leave a CU room for a correlator
// workgroup of the next block (2 waves x 104 VGPRs on a SIMD), where at 144 VGPRs the
// eigen stage kept voltage work off every CU it held for its ~95 us.  The cap spills
// into scratch and costs the kernel ~5 us on its own (384 ch x 36 antennas), but the
// overlapped step gains ~2 %; the Jacobi path (eigen outputs) loses far more under the
// same cap (251 -> 297 us) and keeps the uncapped kernel.
template <int NTHR>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(5))) eig_tri_kernel(EigParams P) {
  extern __shared__ __align__(16) char smem_raw[];
  eig_null_problem<NTHR, false>(P, smem_raw, blockIdx.x >> 1, blockIdx.x & 1);
}

template <typename K>
hipError_t launch_eig_kernel(K kern, const EigParams &p, int nthr, size_t sm, hipStream_t s) {
  if (sm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3(p.C * 2), dim3(nthr), sm, s, p);
  return hipGetLastError();
}

template <int NTHR, bool WARM>
hipError_t launch_eig_t(const EigParams &p, hipStream_t s) {
  const size_t sm = eig_smem_bytes(p.R, p.B, WARM, NTHR / 64);
  if constexpr (!WARM && NTHR == 256) {
    if (!p.full_outputs && !p.shrink && p.R >= 3) return launch_eig_kernel(eig_tri_kernel<NTHR>, p, NTHR, sm, s);
  }
  return launch_eig_kernel(eig_null_kernel<NTHR, WARM>, p, NTHR, sm, s);
}

hipError_t launch_eig_plain(const EigParams &p, int nthr, hipStream_t s) {
  switch (nthr) {
    case 64: return launch_eig_t<64, false>(p, s);
    case 256: return launch_eig_t<256, false>(p, s);
    default: return launch_eig_t<128, false>(p, s);
  }
}

// One launch.  Warm start (fixed K only): a problem whose warm solve is
// rejected runs the plain solver in the same workgroup.
bool eig_is_warm(const EigParams &p) {
  return p.proposal && !p.detect && p.KW >= 2 && p.R >= 2 * p.KW;
}

hipError_t launch_eig(const EigParams &p, hipStream_t s) {
  if (!eig_is_warm(p)) {
    EigParams q = p;
    q.proposal = nullptr;
    return launch_eig_plain(q, p.plain_threads, s);
  }
  return launch_eig_t<256, true>(p, s);
}

// ---- subarrays of uneven size: one persistent launch ----
// Problem q = order[q'] (largest first) is (channel c, pol p, group g): the
// principal submatrix of the (c, p) covariance on the group's antennas (mask
// bits, ascending antenna order), whitened, solved by cyclic Jacobi, K_g
// largest modes removed: outputs eigenvalues (ascending), unwhitened
// normalised eigenvectors and P = I - U_K U_K^H, all in the group's n_g x n_g
// antenna space.  Workgroups take problems from an atomic counter until the
// list is exhausted, so sizes 12 .. 64 share the machine without tail waste.
// One subarray problem q in this workgroup; s_n, s_ant, s_pos: shared
// scratch of the caller (group size, member antennas, antenna -> member).
// With beam operands requested (P.wop), the group's mitigated weights
// W_g' = W_g P_g (W_g: the beam weights on the member antennas) become beam
// rows row0 + g B + b of the (channel, pol)'s operand tiles, zero on the
// other antennas, so the beam kernel forms the subarray beams beside the
// full-array ones.
__device__ __forceinline__ void subarray_problem(const SubParams &P, char *smem_raw, int q, int *s_n, int *s_ant,
                                                 int *s_pos) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long nbl_full = (long)P.R * (P.R + 1) / 2;
  const int g = q % P.G, cp = q / P.G, c = cp >> 1, pol = cp & 1;
  // antenna list of the group: ballot prefix over the mask
  if (tid < 64) {
    const bool on = tid < P.R && P.mask[((long)c * P.G + g) * P.R + tid] != 0;
    const unsigned long long bits = __ballot(on);
    const int pos = __popcll(bits & ((1ull << tid) - 1ull));
    s_pos[tid] = on ? pos : -1;
    if (on) s_ant[pos] = tid;
    if (tid == 0) s_n[0] = __popcll(bits);
  }
  __syncthreads();
  const int n = __builtin_amdgcn_readfirstlane(s_n[0]);
  const long obase = ((long)q) * P.R * P.R;  // output slot of this problem
  if (n == 0) return;
  const int B = P.wop ? P.B : 0;
  EigSmem S = carve(smem_raw, n, B, false, kEigThreads / 64);
  const int lda = jac_ne(n);
  const int nb = n * (n + 1) / 2;
  auto cov = [&](int i, int j) {  // exact A(i, j), i <= j, of the group's antennas
    const int ai = s_ant[i], aj = s_ant[j];
    const long bl = (long)aj * (aj + 1) / 2 + ai;
    long long re = 0, im = 0;
    for (int k = 0; k < P.nsplit; ++k) {
      const int *v = P.vis_part + ((((long)k * P.C + c) * 4 + 3 * pol) * nbl_full + bl) * 2;
      re += v[0];
      im += v[1];
    }
    return make_longlong2(re, im);
  };
  for (int idx = tid; idx < nb; idx += nthr) {  // packed (i <= j) of the subarray
    int j = (int)((sqrtf(8.0f * idx + 1.0f) - 1.0f) * 0.5f);
    while (j * (j + 1) / 2 > idx) --j;
    while ((j + 1) * (j + 2) / 2 <= idx) ++j;
    const int i = idx - j * (j + 1) / 2;
    const longlong2 v = cov(i, j);
    const float2 a = make_float2((float)v.x, (float)v.y);
    S.A[j * lda + i] = a;
    S.A[i * lda + j] = make_float2(a.x, -a.y);
  }
  __syncthreads();
  if (tid < n) S.d[tid] = sqrtf(fmaxf(S.A[tid * lda + tid].x, 1e-10f));
  __syncthreads();
  for (int i = tid; i < n * n; i += nthr) {
    const int col = i / n, row = i - col * n;
    const float den = S.d[row] * S.d[col];
    S.A[col * lda + row].x /= den;
    S.A[col * lda + row].y /= den;
  }
  __syncthreads();
  const int sw = jacobi_lds<kEigThreads>(S.A, S.V, n, P.max_sweeps, P.tol, S.jl());
  sort_eigen(S.A, n, S.perm, S.ev);
  const int K = max(0, min(n, P.k[q]));
  if (K > 0) refine_removed_eigenvalues(S, n, K, 1, cov);
  // unwhitened, normalised eigenvectors (ascending) into A (stride n)
  if (tid < n) {
    const int src = S.perm[tid];
    float nrm = 0.0f;
    for (int i = 0; i < n; ++i) {
      float2 v = S.V[src * S.ldv + i];
      v.x *= S.d[i];
      v.y *= S.d[i];
      S.A[tid * n + i] = v;
      nrm += v.x * v.x + v.y * v.y;
    }
    const float inv = nrm > 0.0f ? rsqrtf(nrm) : 0.0f;
    for (int i = 0; i < n; ++i) {
      S.A[tid * n + i].x *= inv;
      S.A[tid * n + i].y *= inv;
    }
    P.evals[(long)q * P.R + tid] = S.ev[tid];
  }
  __syncthreads();
  float2 *ev_out = reinterpret_cast<float2 *>(P.evecs) + obase;
  float2 *pr_out = reinterpret_cast<float2 *>(P.proj) + obase;
  for (int e = tid; e < n * n; e += nthr) {
    const int col = e / n, row = e - col * n;
    ev_out[e] = S.A[e];
    // P(row, col) = delta - sum_{k >= n-K} u_k(row) conj(u_k(col))
    float2 acc = make_float2(row == col ? 1.0f : 0.0f, 0.0f);
    for (int k = n - K; k < n; ++k) {
      const float2 t = cmulc(S.A[k * n + row], S.A[k * n + col]);
      acc.x -= t.x;
      acc.y -= t.y;
    }
    pr_out[e] = acc;
  }
  if (tid == 0) {
    P.counts[q] = K;
    if (P.sweeps) P.sweeps[q] = sw;
    P.sizes[q] = n;
  }
  if (P.wop) {
    // W_g (rows 0..B-1 of S.W) and G = W_g U_K; W_g' = W_g - G U_K^H (rows B..2B-1)
    for (int idx = tid; idx < B * n; idx += nthr) {
      const int bb = idx / n, i = idx - bb * n;
      const __half2 w = reinterpret_cast<const __half2 *>(P.weights)[((long)cp * B + bb) * P.R + s_ant[i]];
      S.W[idx] = __half22float2(w);
    }
    __syncthreads();
    for (int idx = tid; idx < B * K; idx += nthr) {
      const int bb = idx / K, k = n - K + idx % K;
      float2 acc = make_float2(0.f, 0.f);
      for (int i = 0; i < n; ++i) {
        const float2 t = cmul(S.W[bb * n + i], S.A[k * n + i]);
        acc.x += t.x;
        acc.y += t.y;
      }
      S.G[bb * n + k] = acc;
    }
    __syncthreads();
    for (int idx = tid; idx < B * n; idx += nthr) {
      const int bb = idx / n, i = idx - bb * n;
      float2 acc = S.W[bb * n + i];
      for (int k = n - K; k < n; ++k) {
        const float2 t = cmulc(S.G[bb * n + k], S.A[k * n + i]);
        acc.x -= t.x;
        acc.y -= t.y;
      }
      S.W[(B + bb) * n + i] = acc;
    }
    __syncthreads();
    write_beam_operands(P.wop + (long)cp * P.NMT * P.NK * 2 * 64, P.NMT, P.NK, P.wprec, P.row0 + g * B, B,
                        [&](int e, int a) {
                          const int m = a < P.R ? s_pos[a] : -1;
                          return m >= 0 ? S.W[(B + e) * n + m] : make_float2(0.f, 0.f);
                        });
  }
}

// ---- subarrays of uneven size: one persistent launch ----
// Problem q = order[q'] (largest first) is (channel c, pol p, group g): the
// principal submatrix of the (c, p) covariance on the group's antennas (mask
// bits, ascending antenna order), whitened, solved by cyclic Jacobi, K_g
// largest modes removed: outputs eigenvalues (ascending; the removed ones as
// compensated Rayleigh quotients), unwhitened normalised eigenvectors and
// P = I - U_K U_K^H, all in the group's n_g x n_g antenna space.  Workgroups
// take problems from an atomic counter until the list is exhausted, so sizes
// 12 .. 64 share the machine without tail waste.
__global__ void __launch_bounds__(kEigThreads) subarray_null_kernel(SubParams P) {
  extern __shared__ __align__(16) char smem_raw[];
  __shared__ int s_next, s_n, s_ant[64], s_pos[64];
  for (;;) {
    if (threadIdx.x == 0) s_next = atomicAdd(P.counter, 1);
    __syncthreads();
    const int qi = __builtin_amdgcn_readfirstlane(s_next);  // uniform: scalar registers and branches
    if (qi >= P.nprob) break;  // every wave of the workgroup leaves together
    subarray_problem(P, smem_raw, P.order[qi], &s_n, s_ant, s_pos);
    __syncthreads();  // LDS and the counter slot are reused by the next problem
  }
}

// Every eigenproblem of a block in one persistent launch: the 2C full-array
// (channel, pol) problems (items 0 .. 2C-1, the plain eigen / null stage with
// its [W; W'] operands) and the subarray problems (items 2C + q), taken by
// the workgroups from one largest-first queue.
__global__ void __launch_bounds__(kEigThreads) eig_queue_kernel(EigParams E, SubParams Q, const int *order,
                                                                int ntotal, int *counter) {
  extern __shared__ __align__(16) char smem_raw[];
  __shared__ int s_next, s_n, s_ant[64], s_pos[64];
  for (;;) {
    if (threadIdx.x == 0) s_next = atomicAdd(counter, 1);
    __syncthreads();
    const int qi = __builtin_amdgcn_readfirstlane(s_next);  // uniform: scalar registers and branches
    if (qi >= ntotal) break;
    const int item = order[qi];
    if (item < 2 * E.C)
      eig_null_problem<kEigThreads, false>(E, smem_raw, item >> 1, item & 1);
    else
      subarray_problem(Q, smem_raw, item - 2 * E.C, &s_n, s_ant, s_pos);
    __syncthreads();
  }
}

// The same list with one workgroup per item (item = order[blockIdx.x]): the
// dispatcher starts workgroups in index order as slots free up, so the list
// is still taken largest first, and without the persistent loop the kernel
// needs far fewer registers (values hoisted out of a loop around both item
// bodies stay live across it), which raises its occupancy.
__global__ void __launch_bounds__(kEigThreads) eig_list_kernel(EigParams E, SubParams Q, const int *order) {
  extern __shared__ __align__(16) char smem_raw[];
  __shared__ int s_n, s_ant[64], s_pos[64];
  const int item = order[blockIdx.x];
  if (item < 2 * E.C)
    eig_null_problem<kEigThreads, false>(E, smem_raw, item >> 1, item & 1);
  else
    subarray_problem(Q, smem_raw, item - 2 * E.C, &s_n, s_ant, s_pos);
}

hipError_t launch_eig_queue(const EigParams &e, const SubParams &q, const int *order, int ntotal, int *counter,
                            int num_cus, int max_wgs, hipStream_t s) {
  const size_t sm = std::max(eig_smem_bytes(e.R, e.B, false, kEigThreads / 64),
                             eig_smem_bytes_upto(q.max_n, q.wop ? q.B : 0, false, kEigThreads / 64));
  if (sm > 65536) {
    const hipError_t er = hipFuncSetAttribute(reinterpret_cast<const void *>(eig_queue_kernel),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (er != hipSuccess) return er;
  }
  if (max_wgs <= 0) {  // one workgroup per item
    if (sm > 65536) {
      const hipError_t er = hipFuncSetAttribute(reinterpret_cast<const void *>(eig_list_kernel),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
      if (er != hipSuccess) return er;
    }
    if (ntotal <= 0) return hipSuccess;
    hipLaunchKernelGGL(eig_list_kernel, dim3(ntotal), dim3(kEigThreads), sm, s, e, q, order);
    return hipGetLastError();
  }
  if (hipError_t er = launch_zero(counter, sizeof(int), s)) return er;
  const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / (sm + 1024))));
  const int grid = std::max(1, std::min({ntotal, num_cus * per_cu, max_wgs}));
  hipLaunchKernelGGL(eig_queue_kernel, dim3(grid), dim3(kEigThreads), sm, s, e, q, order, ntotal, counter);
  return hipGetLastError();
}

hipError_t launch_subarray_null(const SubParams &p, int num_cus, int max_wgs, hipStream_t s) {
  // LDS for the largest group; as many workgroups as fit on the machine at once
  const size_t sm = eig_smem_bytes_upto(p.max_n, p.wop ? p.B : 0, false, kEigThreads / 64);
  if (sm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(subarray_null_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  if (hipError_t e = launch_zero(p.counter, sizeof(int), s)) return e;
  const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / (sm + 1024))));
  int grid = std::max(1, std::min(p.nprob, num_cus * per_cu));
  if (max_wgs > 0) grid = std::min(grid, max_wgs);
  hipLaunchKernelGGL(subarray_null_kernel, dim3(grid), dim3(kEigThreads), sm, s, p);
  return hipGetLastError();
}

// ---- standalone batched eigensolver (unit-test entry point) ----
__global__ void __launch_bounds__(kEigThreads)
eigh_batched_kernel(const float2 *Ain, int n, int whiten, int max_sweeps, float tol,
                    float *evals, float2 *evecs, int *sweeps_out) {
  extern __shared__ __align__(16) char smem_raw[];
  EigSmem S = carve(smem_raw, n, 0, false, kEigThreads / 64);
  const int m = blockIdx.x, tid = threadIdx.x, lda = jac_ne(n);
  for (int i = tid; i < n * n; i += blockDim.x) S.A[(i / n) * lda + i % n] = Ain[(long)m * n * n + i];
  __syncthreads();
  if (whiten) {
    if (tid < n) S.d[tid] = sqrtf(fmaxf(S.A[tid * lda + tid].x, 1e-10f));
    __syncthreads();
    for (int i = tid; i < n * n; i += blockDim.x) {
      const int col = i / n, row = i - col * n;
      const float den = S.d[row] * S.d[col];
      S.A[col * lda + row].x /= den;
      S.A[col * lda + row].y /= den;
    }
    __syncthreads();
  }
  const int sw = jacobi_lds<kEigThreads>(S.A, S.V, n, max_sweeps, tol, S.jl());
  sort_eigen(S.A, n, S.perm, S.ev);
  if (tid < n) evals[(long)m * n + tid] = S.ev[tid];
  for (int i = tid; i < n * n; i += blockDim.x) {
    const int k = i / n, r = i % n;
    evecs[(long)m * n * n + i] = S.V[S.perm[k] * S.ldv + r];
  }
  if (tid == 0 && sweeps_out) sweeps_out[m] = sw;
}

// ---- every pol-pair covariance of one block (LambdaGPUPipeline) ----
// LambdaGPUPipeline decomposes each (channel, p, q) covariance of the block,
// batch C*P*P (lambda_gpu_pipeline.hpp:438-439, 494-501): visCorrToDecomp and
// unpack_triangular_baseline_batch (spatial.cuh:673-713) place V[bl(i,j)][p][q]
// at A(i, j), i <= j, and cuSOLVER (fill UPPER) decomposes the Hermitian matrix
// with that upper triangle.  Its samples carry the antenna gains
// (scale_and_convert_to_half in LambdaPipelineIngest, pipeline/common.hpp:
// 505-536), so here A(i, j) = g_ip conj(g_jq) e^{i(phi_i - phi_j)} V, the same
// factor the visibility dump applies.  Outputs: eigenvalues ascending
// [C][P][P][n]; eigenvector k at [C][P][P][k][0..n), scaled so that its
// largest-magnitude component (the first on ties) is real and positive.
__global__ void __launch_bounds__(kEigThreads)
pol_eig_kernel(const int *vis_part, int nsplit, int C, int n, const float2 *gains, const float2 *phase,
               int max_sweeps, float tol, float *evals, float2 *evecs, int *sweeps_out) {
  extern __shared__ __align__(16) char smem_raw[];
  EigSmem S = carve(smem_raw, n, 1, false, kEigThreads / 64);
  const int m = blockIdx.x, tid = threadIdx.x, lda = jac_ne(n);
  const int c = m >> 2, pq = m & 3, p = pq >> 1, q = pq & 1;
  const int nbl = n * (n + 1) / 2;
  for (int idx = tid; idx < nbl; idx += blockDim.x) {
    int j = (int)((sqrtf(8.0f * (float)idx + 1.0f) - 1.0f) * 0.5f);
    while (j * (j + 1) / 2 > idx) --j;
    while ((j + 1) * (j + 2) / 2 <= idx) ++j;
    const int i = idx - j * (j + 1) / 2;
    long long re = 0, im = 0;
    for (int k = 0; k < nsplit; ++k) {
      const int2 v = reinterpret_cast<const int2 *>(vis_part)[(((long)k * C + c) * 4 + pq) * nbl + idx];
      re += v.x;
      im += v.y;
    }
    const float2 gi = gains[((long)c * 2 + p) * n + i], gj = gains[((long)c * 2 + q) * n + j];
    const float2 pi = phase[(long)c * n + i], pj = phase[(long)c * n + j];
    const float2 f = cmul(make_float2(gi.x * gj.x + gi.y * gj.y, gi.y * gj.x - gi.x * gj.y),
                          make_float2(pi.x * pj.x + pi.y * pj.y, pi.y * pj.x - pi.x * pj.y));
    S.A[j * lda + i] = cmul(f, make_float2((float)re, (float)im));
  }
  __syncthreads();
  const int sw = jacobi_lds<kEigThreads>(S.A, S.V, n, max_sweeps, tol, S.jl());
  sort_eigen(S.A, n, S.perm, S.ev);
  float2 *ph = S.W;  // per output column: conj(v_max) / |v_max|
  if (tid < n) {
    const float2 *col = S.V + S.perm[tid] * S.ldv;
    int best = 0;
    float bm = -1.0f;
    for (int r = 0; r < n; ++r) {
      const float mg = col[r].x * col[r].x + col[r].y * col[r].y;
      if (mg > bm) {
        bm = mg;
        best = r;
      }
    }
    const float2 v = col[best];
    const float inv = bm > 0.0f ? rsqrtf(bm) : 0.0f;
    ph[tid] = bm > 0.0f ? make_float2(v.x * inv, -v.y * inv) : make_float2(1.0f, 0.0f);
  }
  __syncthreads();
  if (tid < n) evals[(long)m * n + tid] = S.ev[tid];
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int k = e / n, r = e - k * n;
    evecs[(long)m * n * n + e] = cmul(S.V[S.perm[k] * S.ldv + r], ph[k]);
  }
  if (tid == 0 && sweeps_out) sweeps_out[m] = sw;
}

// ---- projection averaging (LambdaProjectionPipeline) ----
// Per block (lambda_projection_pipeline.hpp:330-431): the same four pol-pair
// covariances as pol_eig_kernel, eigenvectors ascending, then the projector
// onto the K largest, P = U_K U_K^H (cublasCherk, fill UPPER: only i <= j is
// formed), added into the accumulator [C*4][n][n] (column-major, upper).
// Atomic adds: blocks of different buffers may run side by side.
__global__ void __launch_bounds__(kEigThreads)
proj_block_kernel(const int *vis_part, int nsplit, int C, int n, const float2 *gains, const float2 *phase,
                  int K, int max_sweeps, float tol, float2 *acc) {
  extern __shared__ __align__(16) char smem_raw[];
  EigSmem S = carve(smem_raw, n, 1, false, kEigThreads / 64);
  const int m = blockIdx.x, tid = threadIdx.x, lda = jac_ne(n);
  const int c = m >> 2, pq = m & 3, p = pq >> 1, q = pq & 1;
  const int nbl = n * (n + 1) / 2;
  for (int idx = tid; idx < nbl; idx += blockDim.x) {
    int j = (int)((sqrtf(8.0f * (float)idx + 1.0f) - 1.0f) * 0.5f);
    while (j * (j + 1) / 2 > idx) --j;
    while ((j + 1) * (j + 2) / 2 <= idx) ++j;
    const int i = idx - j * (j + 1) / 2;
    long long re = 0, im = 0;
    for (int k = 0; k < nsplit; ++k) {
      const int2 v = reinterpret_cast<const int2 *>(vis_part)[(((long)k * C + c) * 4 + pq) * nbl + idx];
      re += v.x;
      im += v.y;
    }
    const float2 gi = gains[((long)c * 2 + p) * n + i], gj = gains[((long)c * 2 + q) * n + j];
    const float2 pi = phase[(long)c * n + i], pj = phase[(long)c * n + j];
    const float2 f = cmul(make_float2(gi.x * gj.x + gi.y * gj.y, gi.y * gj.x - gi.x * gj.y),
                          make_float2(pi.x * pj.x + pi.y * pj.y, pi.y * pj.x - pi.x * pj.y));
    S.A[j * lda + i] = cmul(f, make_float2((float)re, (float)im));
  }
  __syncthreads();
  jacobi_lds<kEigThreads>(S.A, S.V, n, max_sweeps, tol, S.jl());
  sort_eigen(S.A, n, S.perm, S.ev);
  float2 *out = acc + (long)m * n * n;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int j = e / n, i = e - j * n;
    if (i > j) continue;
    float2 sum = make_float2(0.f, 0.f);
    for (int k = n - K; k < n; ++k) {  // U(i, k) conj(U(j, k))
      const float2 *col = S.V + S.perm[k] * S.ldv;
      const float2 t = cmulc(col[i], col[j]);
      sum.x += t.x;
      sum.y += t.y;
    }
    atomicAdd(&out[e].x, sum.x);
    atomicAdd(&out[e].y, sum.y);
  }
}

// Averaged projector of each (channel, p, q): A = acc / runs (upper triangle),
// its eigenpairs ascending with each eigenvector's largest-magnitude component
// real and positive, and the accumulator cleared (dump_projection,
// lambda_projection_pipeline.hpp:571-645).
__global__ void __launch_bounds__(kEigThreads)
proj_dump_kernel(float2 *acc, int n, float inv_runs, int max_sweeps, float tol, float *evals, float2 *evecs) {
  extern __shared__ __align__(16) char smem_raw[];
  EigSmem S = carve(smem_raw, n, 1, false, kEigThreads / 64);
  const int m = blockIdx.x, tid = threadIdx.x, lda = jac_ne(n);
  float2 *a = acc + (long)m * n * n;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int j = e / n, i = e - j * n;
    const float2 v = a[e];
    a[e] = make_float2(0.f, 0.f);
    if (i <= j) S.A[j * lda + i] = make_float2(v.x * inv_runs, v.y * inv_runs);
  }
  __syncthreads();
  jacobi_lds<kEigThreads>(S.A, S.V, n, max_sweeps, tol, S.jl());
  sort_eigen(S.A, n, S.perm, S.ev);
  float2 *ph = S.W;
  if (tid < n) {
    const float2 *col = S.V + S.perm[tid] * S.ldv;
    int best = 0;
    float bm = -1.0f;
    for (int r = 0; r < n; ++r) {
      const float mg = col[r].x * col[r].x + col[r].y * col[r].y;
      if (mg > bm) {
        bm = mg;
        best = r;
      }
    }
    const float2 v = col[best];
    const float inv = bm > 0.0f ? rsqrtf(bm) : 0.0f;
    ph[tid] = bm > 0.0f ? make_float2(v.x * inv, -v.y * inv) : make_float2(1.0f, 0.0f);
  }
  __syncthreads();
  if (tid < n) evals[(long)m * n + tid] = S.ev[tid];
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int k = e / n, r = e - k * n;
    evecs[(long)m * n * n + e] = cmul(S.V[S.perm[k] * S.ldv + r], ph[k]);
  }
}

hipError_t launch_proj_block(const int *vis_part, int nsplit, int C, int n, const float2 *gains,
                             const float2 *phase, int K, int max_sweeps, float tol, float *acc, hipStream_t s) {
  const size_t sm = eig_smem_bytes(n, 1, false, kEigThreads / 64);
  if (sm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(proj_block_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(proj_block_kernel, dim3(C * 4), dim3(kEigThreads), sm, s, vis_part, nsplit, C, n, gains,
                     phase, K, max_sweeps, tol, reinterpret_cast<float2 *>(acc));
  return hipGetLastError();
}

hipError_t launch_proj_dump(float *acc, int C, int n, int runs, int max_sweeps, float tol, float *evals,
                            float *evecs, hipStream_t s) {
  const size_t sm = eig_smem_bytes(n, 1, false, kEigThreads / 64);
  if (sm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(proj_dump_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(proj_dump_kernel, dim3(C * 4), dim3(kEigThreads), sm, s, reinterpret_cast<float2 *>(acc), n,
                     1.0f / (float)runs, max_sweeps, tol, evals, reinterpret_cast<float2 *>(evecs));
  return hipGetLastError();
}

hipError_t launch_pol_eig(const int *vis_part, int nsplit, int C, int n, const float2 *gains,
                          const float2 *phase, int max_sweeps, float tol, float *evals, float *evecs,
                          int *sweeps, hipStream_t s) {
  const size_t sm = eig_smem_bytes(n, 1, false, kEigThreads / 64);
  if (sm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(pol_eig_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(pol_eig_kernel, dim3(C * 4), dim3(kEigThreads), sm, s, vis_part, nsplit, C, n, gains,
                     phase, max_sweeps, tol, evals, reinterpret_cast<float2 *>(evecs), sweeps);
  return hipGetLastError();
}

// Warm-start solve of whitened Hermitian matrices (test entry): proposal
// [nmat][KW][n], outputs ascending eigenvalues / eigenvectors (columns) and
// the leakage ratio of the K largest modes (accepted when <= tol).
__global__ void __launch_bounds__(kEigThreads)
eigh_warm_kernel(const float2 *Ain, const float2 *prop, int n, int KW, int K, int steps, float tol,
                 float *evals, float2 *evecs, float *ratio) {
  extern __shared__ __align__(16) char smem_raw[];
  EigSmem S = carve(smem_raw, n, 0, true, kEigThreads / 64);
  const int m = blockIdx.x, tid = threadIdx.x, lda = jac_ne(n);
  for (int i = tid; i < n * n; i += blockDim.x) S.A[(i / n) * lda + i % n] = Ain[(long)m * n * n + i];
  for (int e = tid; e < KW * n; e += blockDim.x) S.Y[e] = prop[(long)m * KW * n + e];
  __syncthreads();
  warm_solve(S.A, lda, n, KW, K, steps, tol, S);
  if (tid < n) evals[(long)m * n + tid] = S.ev[tid];
  for (int i = tid; i < n * n; i += blockDim.x) {
    const int k = i / n, r = i % n;
    evecs[(long)m * n * n + i] = S.V[S.perm[k] * S.ldv + r];
  }
  if (tid == 0) ratio[m] = S.red[1];
}

hipError_t launch_eigh_warm(const float *A, const float *prop, int n, int nmat, int KW, int K,
                            int steps, float tol, float *evals, float *evecs, float *ratio,
                            hipStream_t s) {
  const size_t sm = eig_smem_bytes(n, 0, true, kEigThreads / 64);
  if (sm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(eigh_warm_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(eigh_warm_kernel, dim3(nmat), dim3(kEigThreads), sm, s,
                     reinterpret_cast<const float2 *>(A), reinterpret_cast<const float2 *>(prop), n,
                     KW, K, steps, tol, evals, reinterpret_cast<float2 *>(evecs), ratio);
  return hipGetLastError();
}

hipError_t launch_eigh_batched(const float *A, int n, int nmat, int whiten, int max_sweeps,
                               float tol, float *evals, float *evecs, int *sweeps,
                               hipStream_t s) {
  const size_t sm = eig_smem_bytes(n, 0, false, kEigThreads / 64);
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
