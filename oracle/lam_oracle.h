/*
 * lam_oracle.h -- CPU restatement (FP64 / exact integer) of the LAMBDA
 * correlate -> beamform -> RFI-null hot path of jdgsmallwood/cuda-spatial-filtering.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  Nothing in the product links it.
 *
 * Every function cites the reference file:line whose semantics it follows
 * (paths relative to the reference repository root).
 *
 * Layout conventions (all row-major C arrays, complex = [re, im] pairs):
 *   ring samples  int8  [C][NP+2][F][NT][RPP][P][2]   packet_formats.hpp:309-312
 *   ring scales   int16 [C][NP+2][R][P]               packet_formats.hpp:305-306
 *   y (ingested)  int32 [C][T][R][P][2]               T = NP*NT, R = F*RPP
 *   visibilities  int64 [C][R(R+1)/2][P][P][2]        bl(i,j) = j(j+1)/2 + i, i <= j
 *   weights       f64   [C][P][B][R][2]               pipeline/common.hpp:52-55
 *   beams         f64   [C][P][B][T][2]               packet_formats.hpp:346-348
 *   dense matrix  f64   [n][n][2] column-major: element (row i, col j) at j*n + i
 */
#ifndef LAM_ORACLE_H
#define LAM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_shape {
  int C;   /* NR_CHANNELS */
  int F;   /* NR_FPGA_SOURCES */
  int NT;  /* NR_TIME_STEPS_PER_PACKET */
  int R;   /* NR_RECEIVERS (= F * RPP) */
  int P;   /* NR_POLARIZATIONS (2) */
  int RPP; /* NR_RECEIVERS_PER_PACKET */
  int NP;  /* NR_PACKETS_FOR_CORRELATION */
  int B;   /* NR_BEAMS */
} orc_shape;

/* Time-window selection of the two pipelines.
 * ORC_WINDOW_CORR_BEAM: ring slots 0..NP-1 (spatial.cuh:254-264, 337-347).
 * ORC_WINDOW_ADAPTIVE : slot 1 + per-FPGA sample delay (spatial.cuh:1431-1436). */
enum { ORC_WINDOW_CORR_BEAM = 0, ORC_WINDOW_ADAPTIVE = 1 };

/* Ingest: int8 x int16 scale, window/delay selection and stream permutation.
 * spatial.cuh:420-486 (scale), 1420-1506 (delays), 593-649 (permutation).
 * delays may be NULL (all zero); recv_perm/pol_perm may be NULL (identity). */
void orc_ingest(const orc_shape *s, const int8_t *samples, const int16_t *scales,
                int window, const int32_t *delays, const int32_t *recv_perm,
                const int32_t *pol_perm, int32_t *y);

/* Exact integer correlator, packed-triangular output.
 * V[c][bl(i,j)][p][q] = sum_t y[c][t][i][p] * conj(y[c][t][j][q]), i <= j.
 * Conjugation side: tests/test_spatial.cpp:116-125 (67+29j); packing
 * tests/support/assertions.hpp:35-37, src/spatial.cu:139-145. */
void orc_correlate(const orc_shape *s, const int32_t *y, int64_t *vis);

/* Beamformer: beams[c][p][b][t] = sum_r w[c][p][b][r] * y[c][t][r][p]
 * (no conjugation; scripts/adaptive_detection_reference.py:74-77,
 * lambda_corr_beam_only_gpu_pipeline.hpp:479-488). nb = beams in w. */
void orc_beamform(const orc_shape *s, const int32_t *y, const double *w, int nb,
                  double *beams);

/* Dense Hermitian covariance of pol p from packed visibilities, both
 * triangles filled: A(i,j) = V[bl(i,j)][p][p] for i <= j
 * (spatial.cuh:672-713, adaptive pipeline 648-673). */
void orc_covariance(int R, int P, const int64_t *vis_c, int p, double *A);

/* Diagonal whitening: d_k = sqrt(max(Re A_kk, 1e-10)); A_ij /= d_i d_j.
 * spatial.cuh:1065-1093. */
void orc_whiten(int n, double *A, double *d);

/* Cyclic Jacobi eigensolver for a Hermitian matrix (textbook, FP64).
 * Eigenvalues ascending (cusolverDnXsyevBatched contract, adaptive 679-686),
 * eigenvectors as columns of V (column-major). A is destroyed.
 * Returns the number of sweeps used. */
int orc_jacobi_eigh(int n, double *A, double *evals, double *V, int max_sweeps,
                    double tol);

/* Linear-interpolated percentile on sorted values (pipeline/common.hpp:461-478,
 * scripts/adaptive_detection_reference.py:11-25). */
double orc_percentile_linear(const double *sorted, int n, double q);

/* Signal-mode count K = #{lambda > p50 + delta*(p80-p20)/(2*0.8416)}
 * (pipeline/common.hpp:480-495, spatial.cuh:920-945). */
int orc_detect_k(const double *sorted_evals, int n, double delta);

/* u_k = normalize(D v_k), D = diag(d) (spatial.cuh:1095-1125). */
void orc_unwhiten_normalize(int n, double *V, const double *d);

/* P = I - sum_{k=n-K}^{n-1} u_k u_k^H (spatial.cuh:993-1024;
 * scripts/adaptive_detection_reference.py:35-44). */
void orc_null_projection(int n, const double *V, int k, double *P);

/* Shrink: d_k = 1 (k < n-K), lambda_bar/lambda_k (k >= n-K, 0 if lambda_k <= 0);
 * M = V diag(d) V^H (spatial.cuh:952-991, 1036-1063; adaptive 726-756;
 * scripts/adaptive_detection_reference.py:47-63). */
void orc_shrink_projection(int n, const double *V, const double *evals, int k,
                           double *P);

/* W' = W P for nb beam rows of length n (adaptive 262-270, 777-781;
 * scripts/adaptive_detection_reference.py:66-71). W, Wp: [nb][n][2];
 * P column-major [n][n][2]. */
void orc_apply_projection(int nb, int n, const double *W, const double *P,
                          double *Wp);

/* Fine-delay phase table [C][R][N][2]: phi = -2 pi (f_coarse + k_signed/N bw) tau
 * (lambda_gpu_pipeline.hpp:1100-1131). */
void orc_fine_delay_table(int C, int R, int N, const float *delays_ns,
                          double base_freq_hz, double channel_bw_hz,
                          int min_freq_ch, double *table);

/* Full adaptive per-block RFI-null stage for one (channel, pol) problem:
 * covariance -> whiten -> eig -> K -> (unwhiten + I - UU^H | shrink) -> W P.
 * w_cp: [B][R][2] weights of this (c,p); out_w: [2B][R][2] = [W ; W P].
 * evals/evecs receive the pipeline's eigen outputs (adaptive 914-950).
 * fixed_k < 0 selects detection with `delta`. Returns K. */
int orc_rfi_null_cp(int R, int P, int B, const int64_t *vis_c, int p,
                    const double *w_cp, int fixed_k, double delta, int shrink,
                    double *out_w, double *evals, double *evecs);

/* Whole adaptive block (ingest excluded): for every (c,p) run orc_rfi_null_cp
 * then beamform the 2B beams. OpenMP over channels when built with -fopenmp.
 * fixed_k: [C][P] (ignored when detect != 0). Returns 0. */
int orc_adaptive_block(const orc_shape *s, const int32_t *y, const double *w,
                       const int32_t *fixed_k, int detect, double delta,
                       int shrink, int64_t *vis, double *beams2b,
                       double *evals, double *evecs, int32_t *counts);

/* Post-beam spectra of the adaptive pipeline (register_fft_block product):
 * for every (c, p, beam e) and packet o, the unnormalised forward DFT
 *   X[u] = sum_t x[o*nfft + t] exp(-2 pi i u t / nfft)
 * of the packet's nfft samples (cuFFT C2C forward, adaptive 812-814,
 * 1046-1050), drop `drop` bins each side of nfft/2 and put the negative
 * frequencies first (FineChannelCopy / FineChannelSeek flat copies over the
 * frequency-outermost tensor, adaptive 66-76, 816-833, 461-464):
 *   out bin g < keep: u = g + nfft/2 + drop;  g >= keep: u = g - keep,
 *   keep = nfft/2 - drop,
 * and sum |X[u]| over the packets (sum_fft_over_packets, spatial.cuh:1508-1567).
 * beams: [C][P][E][np*nfft][2]; out: [C][P][E][nfft - 2*drop]. Direct O(n^2)
 * DFT in FP64 (no FFT factorisation) so that it checks the FFT independently. */
void orc_beam_spectra(int C, int P, int E, int np, int nfft, int drop,
                      const double *beams, double *out);

/* Bandpass product of LambdaGPUPipeline (lambda_gpu_pipeline.hpp:522-531,
 * 814-816; detect_and_downsample_fft, spatial.cuh:795-835): per row the
 * unnormalised forward DFT over all n samples, then the mean of |X[k]|^2 over
 * each run of d consecutive bins (natural order), NaN values skipped (0 when
 * a run has none).  x: [rows][n][2]; out: [rows][n / d].  Direct O(n^2) DFT in
 * FP64 with exact phase indices (k t mod n). */
void orc_bandpass(int rows, int n, int d, const double *x, double *out);

/* Number of OpenMP threads the oracle would use (1 when built without it). */
int orc_num_threads(void);
/* Threads for the following calls (OpenMP; ignored without it). */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
