// Internal declarations shared by the gfx950 kernels and the C-ABI layer.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lam {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// Beam-weight operand precision (the (hi, lo) operand slots of a k-step):
//   F16_HILO: hi = fp16(w), lo = fp16(w - hi), two MFMAs (about fp32 weights)
//   F16:      hi = fp16(w), one MFMA
//   BF16:     hi = bf16(w) bits, one BF16 MFMA with BF16 samples
constexpr int kWprecF16HiLo = 0, kWprecF16 = 1, kWprecBF16 = 2;

// Weight operand pair for one real weight value.
__device__ __forceinline__ void weight_operands(float v, int wprec, _Float16 &hi, _Float16 &lo) {
  if (wprec == kWprecBF16) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    hi = __builtin_bit_cast(_Float16, (uint16_t)(u >> 16));
    lo = (_Float16)0.0f;
  } else {
    hi = (_Float16)v;
    lo = wprec == kWprecF16HiLo ? (_Float16)(v - (float)hi) : (_Float16)0.0f;
  }
}

// Voltage kernels stage 128-sample chunks through LDS.
constexpr int kChunk = 128;
constexpr int kMaxReceivers = 64;  // 4 antenna groups of 16
constexpr int kMaxSplits = 16;     // time splits per channel (partial visibility buffers)
// Samples per workgroup at most: an int32 partial sum of one split stays exact
// (|Re a Re b + Im a Im b| <= 2 * 128^2 per sample, 2^15 * 2^15 = 2^30 < 2^31).
constexpr int kMaxSamplesPerSplit = 32768;

// Time samples per voltage workgroup (a multiple of kChunk) and the number of
// splits per channel for T samples and C channels.  requested <= 0 picks about
// target_wgs workgroups in total; the result keeps nsplit <= kMaxSplits and every
// split within kMaxSamplesPerSplit.
inline void pick_split(int T, int C, int requested, int &spw, int &nsplit, int target_wgs = 1024) {
  const int Tr = ((T + kChunk - 1) / kChunk) * kChunk;
  spw = requested;
  if (spw <= 0) {
    const int target = (target_wgs + C - 1) / C > 1 ? (target_wgs + C - 1) / C : 1;
    spw = (T + target - 1) / target;
  }
  spw = ((spw + kChunk - 1) / kChunk) * kChunk;
  const int min_spw = (((T + kMaxSplits - 1) / kMaxSplits + kChunk - 1) / kChunk) * kChunk;
  if (spw < min_spw) spw = min_spw;
  if (spw > Tr) spw = Tr;
  if (spw > kMaxSamplesPerSplit) spw = kMaxSamplesPerSplit;
  nsplit = (T + spw - 1) / spw;
}
// Samples per workgroup of a beam-only voltage launch.  It writes nothing per split,
// so it takes more, shorter splits than the correlator (whose split count multiplies
// its partial visibilities): about 12 workgroups per CU, between one chunk and 1024
// samples (2B beams at 384 channels x 8192 samples on 256 CUs: 0.291 ms with the
// correlator's 8192-sample split, 0.252 ms with 1024).
inline int beam_samples_per_wg(int T, int C, int num_cus) {
  const long want = 12L * (num_cus > 0 ? num_cus : 256);
  long spw = ((long)T * C + want - 1) / want;
  spw = ((spw + kChunk - 1) / kChunk) * kChunk;
  if (spw < kChunk) spw = kChunk;
  if (spw > 1024) spw = 1024;
  return (int)spw;
}
#ifndef LAM_EIG_THREADS
#define LAM_EIG_THREADS 256
#endif
constexpr int kEigThreads = LAM_EIG_THREADS;
constexpr int kMaxWarm = 8;        // warm-start vectors per eigenproblem

// Error bits reported through the device flag word.
constexpr int kErrRange = 1;

// Parameters of the voltage kernel (ingest + correlate + beamform).
struct VoltParams {
  const int8_t *samples;   // ring block [C][NP+2][F][NT][RPP][2][2]
  const int16_t *scales;   // [C][NP+2][R][2]
  int C, F, NT, R, RPP, NP, T;
  int window_base;         // 0: corr-beam window (slots 0..NP-1); NT: adaptive
  int nt_log2;             // log2(NT) when NT is a power of two, else -1
  const int *delays;       // [F] per-FPGA sample delays or nullptr
  const int *recv_perm;    // [R*2] or nullptr
  const int *pol_perm;     // [R*2] or nullptr
  int samples_per_wg;      // multiple of kChunk
  int nsplit;              // workgroups per channel
  int xcd_map;             // 1: keep all workgroups of a channel on one XCD
  // correlator output: one int32 partial per time split [nsplit][C][2*p+q][NBL][2]
  int *vis_part;
  // beamformer
  const v8h *wop;          // [C][2][NMT][NK][2][64] weight operands (hi, lo)
  int nbeams_out;          // beams per (channel, pol) written
  int NMT;                 // 32-row operand tiles (16 beams each)
  int NK;                  // K steps of 8 antennas
  __half *beams;           // [C][2][nbeams_out][T][2]
  int wprec;               // beam-weight operand precision (kWprec*)
  int *err;                // error flag word
};

// Parameters of the eigen / RFI-null kernel, one workgroup per (channel, pol).
struct EigParams {
  int C, R, B;
  const int *vis_part;     // per-block partial visibilities [nsplit][C][2*p+q][NBL][2]
  int nsplit;
  const __half *weights;   // [C][2][B][R][2]
  const int *fixed_k;      // [C][2]
  int detect;
  float delta;
  int shrink;
  int max_sweeps;
  float tol;
  int NMT, NK;             // operand tiling of the 2B-beam weights
  v8h *wop;                // out: [C][2][NMT][NK][2][64]
  float *evals;            // out: [C][2][R]
  float *evecs;            // out: [C][2][R][R][2]
  int *counts;             // out: [C][2]
  int *sweeps;             // out: [C][2] or nullptr
  // warm start (learned or previous-block proposal); proposal == nullptr: plain Jacobi
  const float2 *proposal;  // [C*2][prop_stride]: KW vectors of R entries
  int prop_stride;         // float2 per problem
  int KW;                  // warm vectors (even, <= kMaxWarm)
  int refine_sweeps;
  int power;               // 1: one Y <- A Y step before the Householder basis
  float res_tol;           // accept if max_k |A u_k - l_k u_k| <= res_tol * l_max
  float2 *basis_out;       // out: [C*2][KW][R] top-KW whitened eigenvectors, or nullptr
  int *fallback;           // out: [C*2] 1 where the plain solver ran after the check failed
  int *nsig;               // out: [C*2] modes above 4x the noise floor among the K removed, or nullptr
  float *orth;             // out: [C*2] FP64 check max |U_K^H U_K - I| of the removed (whitened) basis
  int plain_threads;       // workgroup width of the plain solver (64, 128, 256)
  int wprec;               // precision of the [W; W'] beam operands (kWprec*)
  int whiten;              // 1: whiten the covariance first (adaptive); 0: raw (pulsar-fold RFI branch)
  int full_outputs;        // 1: every eigenpair is written (eigen outputs requested); 0: only what the
                           // null needs (the warm start then forms no complement basis)
};

// Eigen stage of subarrays (uneven sizes), one persistent launch.
struct SubParams {
  int C, R, G;             // channels, receivers, groups per channel
  const int *vis_part;     // [nsplit][C][2*p+q][NBL][2]
  int nsplit;
  const uint8_t *mask;     // [C][G][R] antenna membership
  const int *k;            // [C*2*G] modes to remove
  const int *order;        // [nprob] problem ids, largest first
  int nprob;               // C * 2 * G
  int max_n;               // largest group
  int max_sweeps;
  float tol;
  int *counter;            // work-queue counter (cleared by the launcher)
  float *evals;            // out: [nprob][R] (first n_g valid)
  float *evecs;            // out: [nprob][R*R][2] (n_g x n_g, columns)
  float *proj;             // out: [nprob][R*R][2] P = I - U_K U_K^H (n_g x n_g)
  int *counts, *sweeps, *sizes;  // out: [nprob]
  // subarray beams (wop == nullptr: none): W_g' = W_g P_g as beam rows
  // row0 + g * B + b of the (channel, pol) operand tiles [NMT][NK][2][64]
  const __half *weights;   // [C][2][B][R][2]
  int B, NMT, NK, row0, wprec;
  v8h *wop;
};

// Synthetic voltage generator (benchmark / test input).
constexpr int kGenMaxSources = 8;
struct GenParams {
  int C, F, NT, R, RPP, NP;
  int S;                   // tone slots per channel
  unsigned long long seed;
  int block_index;
  float sigma;             // noise per real component
  const float *tones;      // [C][S][6]: amplitude, cycles/sample, phase, pol-0 gain, pol-1 gain (re, im)
  const float2 *steer;     // [C][S][R]
  const float *gain;       // [C] requantisation gain
  int8_t *samples;         // ring [C][NP+2][F][NT][RPP][2][2]
  int16_t *scales;         // [C][NP+2][R][2]
};

// Features of the learned warm start, one row per (channel, pol).
struct FeatParams {
  const int *vis_part;     // [nsplit][C][2*p+q][NBL][2]
  int nsplit, C, n, KW, Dp;
  const float2 *basis;     // [C][2][KW][n] previous block's top-KW whitened eigenvectors or nullptr
  uint16_t *X;             // out: [C*2][Dp] BF16
};

// Launchers (defined in the .hip files).
hipError_t launch_features(const FeatParams &p, hipStream_t s);
// Post-beam spectra: out [C][2][nbeams][NT - 10] (NT = 64) from FP16 beams [C][2][nbeams][T].
// Post-beam spectra: 64-point DFT magnitudes per packet, 54 kept bins, summed
// over packets.  table: spectra_table_bytes() of device memory filled from
// spectra_table_host().
// Projection averaging (LambdaProjectionPipeline): per block add the
// projector onto the K largest eigenvectors of every pol-pair covariance into
// acc [C*4][n][n][2] (upper, column-major); the dump decomposes acc / runs into
// evals [C][2][2][n] / evecs [C][2][2][n][n][2] and clears acc.
hipError_t launch_proj_block(const int *vis_part, int nsplit, int C, int n, const float2 *gains,
                             const float2 *phase, int K, int max_sweeps, float tol, float *acc, hipStream_t s);
hipError_t launch_proj_dump(float *acc, int C, int n, int runs, int max_sweeps, float tol, float *evals,
                            float *evecs, hipStream_t s);
// LambdaGPU bandpass: N-point FFT per beam row (N = block length, power of
// two <= 8192), mean |X|^2 over runs of D bins -> out [rows][N / D].
hipError_t launch_bandpass(const __half *beams, int rows, int N, int D, float *out, hipStream_t s);
// Eigendecomposition of all four pol-pair covariances of one block
// (LambdaGPUPipeline): evals [C][2][2][n], evecs [C][2][2][n][n][2].
hipError_t launch_pol_eig(const int *vis_part, int nsplit, int C, int n, const float2 *gains,
                          const float2 *phase, int max_sweeps, float tol, float *evals, float *evecs,
                          int *sweeps, hipStream_t s);
size_t spectra_table_bytes();
void spectra_table_host(uint16_t *dst);
hipError_t launch_spectra(const __half *beams, const void *table, int C, int nbeams, int T, int NT, float *out,
                          hipStream_t s);
struct Skip {
  const uint16_t *X = nullptr;  // BF16 rows holding the skip source
  int ld = 0, off = 0, len = 0;
  float scale = 0.0f;
};
// Inference forward of the 2-layer MLP: hidden layer relu(X W1^T + b1) as BF16
// in Hbf (M * H), output Y = Hbf W2^T + b2 + skip (M * N FP32).
hipError_t launch_mlp_infer(const uint16_t *X, int M, int K, const uint16_t *W1, const float *b1, int H,
                            const uint16_t *W2, const float *b2, int N, uint16_t *Hbf, float *Y, const Skip &sk,
                            hipStream_t s);
hipError_t launch_dense_fwd(const uint16_t *X, int M, int K, const uint16_t *Wt, const float *b,
                            int N, float *Z, uint16_t *Hbf, const Skip &sk, hipStream_t s);
hipError_t launch_gemm_bwd(bool at, const float *A, int lda, const uint16_t *B, int ldb, float *C,
                           int ldc, int I, int J, int K, const float *mask, int ldm, hipStream_t s);
hipError_t launch_colsum(const float *D, int M, int N, float *out, hipStream_t s);
hipError_t launch_loss(const float *Y, int Dyp, const float2 *U, const int *K, int M, int n,
                       int KW, float *dY, float *loss_acc, hipStream_t s);
hipError_t launch_adam(float *prm, const float *g, float *m1, float *m2, long P, float lr, float b1,
                       float b2, float eps, int t, long w1_off, long w1_len, uint16_t *w1bf,
                       long w2_off, long w2_len, uint16_t *w2bf, hipStream_t s);
hipError_t launch_to_bf16(const float *src, long n, uint16_t *dst, hipStream_t s);
hipError_t launch_gather(const int *idx, int nb, const uint16_t *DX, int Dp, const float2 *DU,
                         int nu, const int *DK, uint16_t *X, float2 *U, int *K, hipStream_t s);
hipError_t launch_synth(const GenParams &g, hipStream_t s);
hipError_t launch_volt(const VoltParams &p, bool do_corr, bool do_beam, hipStream_t s);
hipError_t launch_eig(const EigParams &p, hipStream_t s);
// num_cus: compute units to fill; max_wgs > 0 caps the persistent grid.
hipError_t launch_subarray_null(const SubParams &p, int num_cus, int max_wgs, hipStream_t s);
// Full-array and subarray problems of a block from one largest-first queue
// order[ntotal] (item < 2C: (channel, pol) = (item / 2, item % 2); else
// subarray problem item - 2C).
hipError_t launch_eig_queue(const EigParams &e, const SubParams &q, const int *order, int ntotal, int *counter,
                            int num_cus, int max_wgs, hipStream_t s);
bool eig_is_warm(const EigParams &p);
hipError_t launch_prep_weights(const __half *w, const float2 *gains, const float2 *phase,
                               int C, int B, int R, int NMT, int NK, int wprec, v8h *wop,
                               hipStream_t s);
hipError_t launch_accumulate_vis(const int *vis_part, int nsplit, long long count, int nbl, long long *acc,
                                 bool atomic, hipStream_t s);
hipError_t launch_dump_vis(long long *acc, int C, int R, const float2 *gains,
                           const float2 *phase, float *out_f32, long long *out_i64,
                           hipStream_t s);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_eigh_warm(const float *A, const float *prop, int n, int nmat, int KW, int K,
                            int steps, float tol, float *evals, float *evecs, float *ratio,
                            hipStream_t s);
hipError_t launch_eigh_batched(const float *A, int n, int nmat, int whiten, int max_sweeps,
                               float tol, float *evals, float *evecs, int *sweeps,
                               hipStream_t s);

}  // namespace lam
