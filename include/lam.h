/*
 * lam.h -- C ABI of the MI355X (gfx950) LAMBDA correlate -> beamform -> RFI-null
 * hot path.  Plain pointers, sizes and status codes; no C++ or torch types.
 *
 * The ABI replaces the device half of the reference's two GPU pipelines
 * (jdgsmallwood/cuda-spatial-filtering, paths relative to its root):
 *   LambdaCorrBeamOnlyGPUPipeline        include/spatial/pipeline/lambda_corr_beam_only_gpu_pipeline.hpp
 *   LambdaAdaptiveBeamformedSpectraPipeline
 *                                        include/spatial/pipeline/lambda_adaptive_beamformed_spectra_pipeline.hpp
 * behind the GPUPipeline seam (include/spatial/pipeline_base.hpp:9-40).  The
 * C++ classes of the same names (cuda-spatial-filtering_amd/include/spatial/)
 * are thin host wrappers over these entry points.
 *
 * Layouts (row-major; "complex" = interleaved re, im):
 *   samples      int8     [C][NP+2][F][NT][RPP][P][2]   packet_formats.hpp:309-312
 *   scales       int16    [C][NP+2][R][P]               packet_formats.hpp:305-306
 *   weights      float16  [C][P][B][R][2]               pipeline/common.hpp:52-55
 *   gains        float32  [C][P][R][2]                  packet_formats.hpp:384-387
 *   beams        float16  [C][P][Bout][NP*NT][2]        packet_formats.hpp:346-348
 *                (Bout = B in corr-beam mode, 2B in adaptive mode: beams
 *                 0..B-1 original, B..2B-1 RFI-mitigated, adaptive 773-806;
 *                 with G subarray groups (lam_set_subarrays) (2 + G)B: beams
 *                 2B + gB .. 2B + gB + B-1 mitigated on group g's antennas)
 *   visibilities float32  [C][R(R+1)/2][P][P][2]        packet_formats.hpp:354-356
 *                V[bl(i,j)][p][q] = sum_t x_ip conj(x_jq), bl(i,j) = j(j+1)/2 + i, i <= j
 *   eigenvalues  float32  [C][P][R]     ascending       adaptive 41-45, 914-950
 *   eigenvectors float32  [C][P][R][R][2] column-major (eigenvector k at [k][0..R))
 *   counts       int32    [C][P]                        adaptive 709-715
 *   pol-pair eigen (LAM_MODE_LAMBDA_GPU)                lambda_gpu_pipeline.hpp:44-49
 *                eigenvalues  float32 [C][P][P][R] ascending
 *                eigenvectors float32 [C][P][P][R][R][2], eigenvector k at [k][0..R),
 *                largest-magnitude component real and positive
 *   spectra      float32  [C][P][2B][NT-10]             adaptive 808-848, FFTOutputType
 *                sum over the block's packets of |64-point FFT| of each beam
 *                packet, 5 edge bins dropped each side, DC centred
 *
 * Threading: a context is used from one host thread (the reference's
 * pipeline_feeder, spatial.hpp:1275-1299).  Host callbacks run on the HIP
 * runtime's callback thread, as cudaLaunchHostFunc callbacks do in the
 * reference (pipeline/common.hpp:387-459).
 */
#ifndef LAM_H
#define LAM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LAM_ABI_VERSION 1

/* Status codes.  The reference exits the process on CUDA errors
 * (logging.hpp:11-27); this ABI returns them instead. */
enum {
  LAM_OK = 0,
  LAM_ERR_ARG = 1,         /* invalid argument or shape (reference: std::runtime_error) */
  LAM_ERR_HIP = 2,         /* HIP runtime failure */
  LAM_ERR_RANGE = 3,       /* a scaled sample left the int8 range of the exact correlator */
  LAM_ERR_STATE = 4,       /* call out of order (reference: std::logic_error, common.hpp:513-515) */
  LAM_ERR_UNSUPPORTED = 5, /* shape outside this build's limits */
  LAM_ERR_NO_DEVICE = 6    /* no gfx950 device */
};

/* Pipeline variants. */
enum {
  LAM_MODE_CORR_BEAM = 0, /* LambdaCorrBeamOnlyGPUPipeline */
  LAM_MODE_ADAPTIVE = 1,  /* LambdaAdaptiveBeamformedSpectraPipeline (corr + eig + null + 2B beams) */
  /* LambdaGPUPipeline's correlate / beamform / eigen part (lambda_gpu_pipeline.hpp:445-608):
   * B beams and accumulated visibilities as in corr-beam mode, on the delay-aligned
   * window of the adaptive pipeline (slot 1 + per-FPGA delay, stream permutation),
   * plus the eigendecomposition of all four pol-pair covariances of every block
   * (layout "pol-pair eigen" below); with lam_set_bandpass, also its bandpass (the
   * block-length FFT of every beam row, mean |X|^2 over runs of bins). */
  LAM_MODE_LAMBDA_GPU = 2,
  /* LambdaProjectionPipeline (lambda_projection_pipeline.hpp:307-645): per block the
   * four pol-pair covariances (ring slots 0..NP-1, gains applied) are decomposed and
   * the projector onto their K largest eigenvectors is accumulated; a projection
   * dump decomposes the average.  No beams and no visibility product. */
  LAM_MODE_PROJECTION = 3
};

/* Runtime form of LambdaConfig (packet_formats.hpp:265-407). */
typedef struct lam_shape {
  int32_t nr_channels;                /* C */
  int32_t nr_fpga_sources;            /* F */
  int32_t nr_time_steps_per_packet;   /* NT */
  int32_t nr_receivers;               /* R = F * RPP, <= 64 */
  int32_t nr_polarizations;           /* P, must be 2 */
  int32_t nr_receivers_per_packet;    /* RPP */
  int32_t nr_packets_for_correlation; /* NP */
  int32_t nr_beams;                   /* B */
  int32_t nr_padded_receivers;        /* reference TCC padding; informational */
  int32_t nr_blocks_to_accumulate;    /* NR_CORRELATED_BLOCKS_TO_ACCUMULATE */
} lam_shape;

typedef struct lam_ctx lam_ctx;
typedef void (*lam_host_fn)(void *user);

/* Create a pipeline context on the current HIP device with `num_buffers`
 * round-robin buffers/streams (lambda_corr_beam_only_gpu_pipeline.hpp:385-470;
 * adaptive 1014-1210).  Weights start at zero, gains at 1+0j, delays 0. */
int lam_create(const lam_shape *shape, int mode, int num_buffers, lam_ctx **out);
int lam_destroy(lam_ctx *ctx);
/* Last error text of this context ("" if none); valid until the next call. */
const char *lam_last_error(const lam_ctx *ctx);
int lam_abi_version(void);

/* Beam weights complex<half>[C][P][B][R] (BeamWeightsT, pipeline/common.hpp:52-55);
 * copied (lambda_corr_beam_only_gpu_pipeline.hpp:491-493). */
int lam_set_weights(lam_ctx *ctx, const uint16_t *weights_half);
/* Antenna gains complex<float>[C][P][R] (GPUPipeline::set_antenna_gains,
 * pipeline_base.hpp:17-19).  Applied as g_i conj(g_j) to visibilities and
 * folded into beam weights. */
int lam_set_gains(lam_ctx *ctx, const float *gains);
/* Per-FPGA integer sample delays, F ints in [-NT, NT]
 * (set_subpacket_delays, adaptive 1216-1220; apply_delays spatial.cuh:1420-1506). */
int lam_set_subpacket_delays(lam_ctx *ctx, const int32_t *delays);
/* Per-receiver fine delays in ns (set_fine_delays,
 * lambda_gpu_pipeline.hpp:1100-1131), applied as the channel-centre phase
 * exp(-2 pi i f_c tau) folded into the beam weights and exp(i(phi_i - phi_j))
 * on visibilities.  delays_ns == NULL clears them. */
int lam_set_fine_delays(lam_ctx *ctx, const float *delays_ns, double base_freq_hz,
                        double channel_bw_hz, int32_t min_freq_ch);
/* Stream permutation (set_stream_permutation, adaptive 549-561):
 * n must equal F*RPP*P, entries indexed by canonical (receiver, pol); recv -1
 * gives zeros.  LAM_ERR_ARG on size mismatch. */
int lam_set_stream_permutation(lam_ctx *ctx, const int32_t *recv_perm,
                               const int32_t *pol_perm, size_t n);
/* Fixed number of nulled modes per (channel, pol), int32[C][P] (adaptive 1157-1167). */
int lam_set_fixed_k(lam_ctx *ctx, const int32_t *k);
/* Null/shrink and detection switches (adaptive ctor 1014-1020, 687-757). */
int lam_set_nulling(lam_ctx *ctx, int detect, float delta, int shrink);

/* Whitening of each covariance before the eigen solve (adaptive mode; on by
 * default, spatial.cuh:1065-1093).  0 is the pulsar-fold pipeline's RFI branch
 * (lambda_pulsar_fold_pipeline.hpp:481-640): the raw XX / YY covariance is
 * decomposed, P = I - U_K U_K^H with its K largest eigenvectors, W' = W P, and
 * the eigen outputs are the raw covariance's.  Plain solver only, and no
 * subarrays (their eigen stage always whitens): LAM_ERR_STATE otherwise, in
 * either order. */
int lam_set_whitening(lam_ctx *ctx, int enable);

/* Post-beam spectra of the adaptive pipeline (register_fft_block,
 * lambda_adaptive_beamformed_spectra_pipeline.hpp:808-848, 1173-1178): when
 * enabled, each block also computes, per (channel, pol, output beam), the
 * 64-point forward FFT of every packet's time series, keeps the 54 bins
 * that FineChannelCopy keeps (DC centred), and sums their magnitudes over
 * the block's packets into the spectra layout above.  Needs NT = 64.  Off
 * by default (it reads the whole beam output once more per block).
 * LAM_ERR_ARG in corr-beam mode or for another NT. */
int lam_set_spectra(lam_ctx *ctx, int enable);

/* Bandpass product of the LambdaGPU mode (lambda_gpu_pipeline.hpp:522-531,
 * 814-816; detect_and_downsample_fft, spatial.cuh:795-835): per (channel,
 * pol, beam) the forward FFT over the whole block (NP * NT points), then the
 * mean of |X|^2 over each run of `downsample_factor` consecutive bins (natural
 * FFT order, NaN skipped) -> float32 [C][P][B][NP*NT / downsample_factor],
 * returned through lam_block_io.spectra_dst.  0 turns it off (the default; a
 * no-op in the other modes).  LAM_ERR_ARG for a factor > 0 outside the
 * LambdaGPU mode, for a block length that is not a power of two up to 8192,
 * or for a factor that does not divide it. */
int lam_set_bandpass(lam_ctx *ctx, int downsample_factor);
/* Eigensolver controls: maximum Jacobi sweeps and convergence tolerance on
 * the relative off-diagonal Frobenius norm. */
int lam_set_eig_params(lam_ctx *ctx, int max_sweeps, float tol);

/* One block's inputs and output landing pointers (the ProcessorState ->
 * GPUPipeline -> Output seam, pipeline/common.hpp:505-537,
 * lambda_corr_beam_only_gpu_pipeline.hpp:259-284, adaptive 850-951). */
typedef struct lam_block_io {
  const void *samples;     /* int8 ring block */
  const void *scales;      /* int16 scales */
  int32_t inputs_on_device;/* 1: device pointers, no H2D copy */
  lam_host_fn release_fn;  /* called once the host input is no longer read (may be NULL) */
  void *release_user;
  void *beams_dst;         /* host or device landing pointer for beams (NULL: skip) */
  lam_host_fn beams_done_fn;
  void *beams_done_user;
  float *eigenvalues_dst;  /* adaptive or LAMBDA_GPU mode (NULL: skip).  In the adaptive
                            * mode a block with neither eigen landing pointer forms only
                            * what its null needs: no eigen outputs are written, the
                            * plain solve finds the eigenvalues K needs and only the
                            * removed modes' vectors (tridiagonal reduction instead of
                            * Jacobi, except in shrink mode), and the warm starts form no
                            * complement basis; the 2B beams and counts agree with a block
                            * that requests the outputs to the beams' tolerance. */
  float *eigenvectors_dst;
  int32_t *counts_dst;
  lam_host_fn eigen_done_fn;
  void *eigen_done_user;
  float *spectra_dst;      /* adaptive spectra (lam_set_spectra) or LambdaGPU bandpass
                            * (lam_set_bandpass); NULL: skip */
  lam_host_fn spectra_done_fn;
  void *spectra_done_user;
} lam_block_io;

/* Enqueue one block on the next buffer's stream (execute_pipeline,
 * pipeline_base.hpp:28-29).  Returns after enqueueing.  When io->release_fn
 * is set it is called exactly once per call, also when the call fails: from
 * the stream after the H2D copy (or after the block, for device inputs), or,
 * if the call fails before that was enqueued, directly before returning.
 * A block whose scaled samples left the int8 range (LAM_ERR_RANGE) is
 * reported by the first lam_execute, lam_dump_* or lam_sync call after it
 * has completed: that call fails with LAM_ERR_RANGE, names the block in
 * lam_last_error, and runs no new block (lam_sync: waits first). */
int lam_execute(lam_ctx *ctx, const lam_block_io *io);

/* Enqueue a copy of the visibility accumulator into dst (float32 layout
 * above, gains/phases applied) and reset it (dump_visibilities,
 * lambda_corr_beam_only_gpu_pipeline.hpp:660-701).  dst may be host or
 * device memory; done_fn runs when the copy has landed. */
int lam_dump_visibilities(lam_ctx *ctx, float *dst, lam_host_fn done_fn, void *user);
/* Exact form of the same: int64 [C][R(R+1)/2][P][P][2] raw integer sums
 * (no gains or phases), then reset. */
int lam_dump_visibilities_i64(lam_ctx *ctx, int64_t *dst, lam_host_fn done_fn, void *user);
/* Blocks accumulated since the last dump. */
int lam_blocks_accumulated(const lam_ctx *ctx);

/* Projection mode: K = nr_signal_eigenvectors (1..R) and the number of blocks
 * averaged per dump (the NR_SIGNAL_EIGENVECTORS / NR_RUNS_TO_AVERAGE template
 * arguments; defaults 1, 1).  lam_projection_runs: blocks accumulated since the
 * last dump; a caller dumps when it reaches nr_runs_to_average, as
 * execute_pipeline does (lambda_projection_pipeline.hpp:432-438). */
int lam_set_projection(lam_ctx *ctx, int nr_signal_eigenvectors, int nr_runs_to_average);
int lam_projection_runs(const lam_ctx *ctx);
/* Enqueue dump_projection (lambda_projection_pipeline.hpp:571-645): the average of
 * the accumulated projectors U_K U_K^H (upper triangle, as cublasCherk fills it) is
 * decomposed; eigenvalues [C][P][P][R] ascending and eigenvectors
 * [C][P][P][R][R][2] (eigenvector k at [k][0..R), largest-magnitude component
 * real and positive) land in evals_dst / evecs_dst (host or device, either may
 * be NULL); the accumulator is cleared.  LAM_ERR_STATE when no block was
 * accumulated since the last dump. */
int lam_dump_projection(lam_ctx *ctx, float *evals_dst, float *evecs_dst, lam_host_fn done_fn, void *user);

/* Wait for all enqueued work; reports LAM_ERR_RANGE if a block since the
 * last report had samples scaled out of the int8 range. */
int lam_sync(lam_ctx *ctx);

/* Device stream of the buffer that the next lam_execute will use. */
void *lam_next_stream(lam_ctx *ctx);

/* hipGraph capture of the per-block kernel sequence: 1 = use graphs
 * (default).  SPATIAL_DISABLE_CUDA_GRAPH in the environment (any value), as
 * in the reference (lambda_corr_beam_only_gpu_pipeline.hpp:561), or
 * LAM_DISABLE_GRAPH=1 turn it off at lam_create.  A failed capture or
 * instantiation prints one warning to stderr and the context runs eagerly
 * from then on (lambda_corr_beam_only_gpu_pipeline.hpp:566-577). */
int lam_set_graphs(lam_ctx *ctx, int enable);
/* Number of graph replays (hipGraphLaunch) so far; stays 0 when every block
 * ran eagerly.  -1 for a NULL context. */
long long lam_graph_replays(const lam_ctx *ctx);

/* Per-kernel timing: when enabled, HIP events bracket every kernel of each
 * block on its stream; lam_kernel_times returns the accumulated milliseconds
 * and launch counts per kernel id (see LAM_K_* below) and resets them.
 * LAM_K_EIGFB keeps its number but records nothing: a rejected warm solve is
 * re-solved inside the warm launch. */
enum {
  LAM_K_CORRBEAM = 0, LAM_K_CORR = 1, LAM_K_EIG = 2, LAM_K_BEAM2B = 3,
  LAM_K_FEAT = 4, LAM_K_MLP = 5, LAM_K_EIGFB = 6, LAM_K_SUBARRAY = 7, LAM_K_SPECTRA = 8,
  LAM_K_POLEIG = 9, LAM_K_BANDPASS = 10, LAM_K_COUNT = 11
};
int lam_set_profiling(lam_ctx *ctx, int enable);
int lam_kernel_times(lam_ctx *ctx, double *ms, int32_t *launches);

/* Tunables picked by the autotuner: time samples per workgroup of the
 * correlating voltage kernels, rounded up to a multiple of 128; 0 = automatic
 * (about one correlator workgroup per CU in the adaptive mode, about 1024
 * workgroups in the others).  The adaptive mode's 2B-beam launch, which writes
 * nothing per split, always takes its own finer split (about 12 workgroups per CU,
 * at most 1024 samples each). */
int lam_set_tuning(lam_ctx *ctx, int samples_per_wg);
/* Precision of the beamformer's weight operands (no counterpart in the
 * reference, whose ccglib GEMM takes FP16 weights):
 *   LAM_WPREC_F16_HILO  weights split into FP16 hi + lo, two FP16 MFMAs (default;
 *                       about FP32 weights, incl. the projected W' = W P)
 *   LAM_WPREC_F16       FP16 weights, one FP16 MFMA
 *   LAM_WPREC_BF16      BF16 weights and samples, one BF16 MFMA
 * The int8 samples are exact in every case; accumulation is FP32. */
enum { LAM_WPREC_F16_HILO = 0, LAM_WPREC_F16 = 1, LAM_WPREC_BF16 = 2 };
int lam_set_weight_precision(lam_ctx *ctx, int precision);

/* Workgroup width (64, 128 or 256 threads: one to four wavefronts per
 * matrix) of the plain eigen launch (Jacobi or the tridiagonal solve); 256 by
 * default.  Results are identical bit for bit across launches of the same width. */
int lam_set_eig_tuning(lam_ctx *ctx, int plain_threads);

/* Autotuner (no counterpart in the reference): times this context's kernels
 * on a sequence of nblocks device-resident blocks (samples[i], scales[i]),
 * `iters` timed passes after one untimed pass, and keeps the fastest
 *   - time samples per voltage workgroup: automatic, 1024, 2048, 3072, 4096,
 *     8192, judged on the whole block sequence (more splits speed the voltage
 *     kernels but multiply the partial visibilities the eigen stage reads);
 *   - plain eigen solve workgroup width: 64, 128, 256 (adaptive mode);
 *   with several buffers both are judged on streaming time per block (blocks
 *   overlapping, at least 16 timed passes), with one buffer on kernel time;
 *   - warm-start refinement steps 1..3 when LAM_EIG_LEARNED / PREVIOUS is set
 *     (rejected problems fall back to the plain solver inside the timing, so a
 *     too-short refinement loses on time).
 * Each choice measures its candidates in three interleaved rounds (order
 * rotated per round) after a warm-up of about 48 blocks, judges each by its
 * median, and keeps the setting in force before the call unless the best
 * median beats it by more than that setting's own spread over the rounds.
 * The eigen stage is timed as the context's most recent lam_execute block ran it
 * (with or without eigen outputs: Jacobi or the tridiagonal solve; a context that
 * has executed no block yet is tuned without eigen outputs).
 * The blocks do not reach the visibility accumulator, and training capture
 * (lam_learn_capture) is suspended while tuning.  result[3] (may be NULL)
 * = {samples_per_wg, plain_threads, refine_steps}. */
int lam_autotune(lam_ctx *ctx, const void *const *samples, const void *const *scales, int nblocks,
                 int iters, int32_t *result);

/* Pinned host memory for block buffers and output landing areas
 * (cudaHostAlloc in LambdaFinalPacketData, packet_formats.hpp:131-142, and
 * SingleHostMemoryOutput, output.hpp:146-154).  Returns NULL on failure. */
void *lam_host_alloc(size_t bytes);
void lam_host_free(void *p);

/* Device memory for callers without their own HIP runtime (the Python
 * binding, tests and benchmarks): plain hipMalloc/hipFree/hipMemcpy on the
 * current device.  lam_memcpy is synchronous and accepts host or device
 * pointers on either side. */
void *lam_device_alloc(size_t bytes);
void lam_device_free(void *p);
int lam_memcpy(void *dst, const void *src, size_t bytes);
/* Number of visible gfx950 devices (0 when none, e.g. on a build host). */
int lam_device_count(void);

/* ---- eigen-stage method and learned warm start (adaptive mode) ----
 * No counterpart in the reference, which always calls cusolverDnXsyevBatched
 * (lambda_adaptive_beamformed_spectra_pipeline.hpp:679-686).
 *
 * LAM_EIG_PLAIN     cyclic Jacobi from the identity (the default).
 * LAM_EIG_LEARNED   a 2-layer MLP (lam_learn_*) reads the whitened covariance
 *                   upper triangle and the previous block's top eigenvectors
 *                   and proposes a warm-start subspace of `warm_vectors`
 *                   vectors; `refine_sweeps` subspace-iteration steps
 *                   (Y <- orth(A Y)) and a Rayleigh-Ritz Jacobi solve of the
 *                   projected warm_vectors x warm_vectors matrix refine it;
 *                   the result is accepted only if the interference the K
 *                   removed modes would leak, sum_k |A u_k - l_k u_k|^2 / l_k
 *                   x (l_k - mu) / l_k, is at most residual_tol x mu (mu: mean
 *                   of the remaining eigenvalues); otherwise the plain solver
 *                   runs on that matrix.  Needs fixed K (detection needs the
 *                   whole spectrum and always runs the plain solver).
 * LAM_EIG_PREVIOUS  the same refinement with the previous block's vectors as
 *                   the proposal (no network).
 * The eigenpairs outside the warm subspace are an orthonormal complement with
 * its Rayleigh quotients.  `power` is reserved (pass 0).
 * With LEARNED / PREVIOUS, consecutive blocks are ordered (each block's warm
 * start depends on its predecessor). */
enum { LAM_EIG_PLAIN = 0, LAM_EIG_LEARNED = 1, LAM_EIG_PREVIOUS = 2 };
int lam_set_eig_method(lam_ctx *ctx, int method, int refine_sweeps, int power, float residual_tol);
/* Per (channel, pol) of the most recent block: Jacobi sweeps run (refinement
 * plus fallback), whether the plain fallback ran, and the FP64 orthogonality
 * check max |u_a^H u_b - delta_ab| of the K removed whitened eigenvectors.
 * Any pointer may be NULL.  Synchronous. */
int lam_eig_stats(lam_ctx *ctx, int32_t *sweeps, int32_t *fallback, float *orthogonality);

/* Subarrays (no counterpart in the reference): ngroups antenna groups per
 * channel, mask[C][ngroups][R] (non-zero = member), k[C][2][ngroups] modes to
 * remove.  Every block then also solves all C x 2 x ngroups eigenproblems of
 * uneven size (the principal submatrices of each pol's covariance on the
 * group's antennas, whitened, plain Jacobi, the removed eigenvalues refined
 * as compensated Rayleigh quotients), forms each group's null projector
 * P_g = I - U_K U_K^H in its own antenna space and its mitigated beams with
 * the weights W_g P_g (W_g: the beam weights on the group's antennas) as
 * output beams 2B + gB .. (see the beams layout; (2 + ngroups) x B <= 64).
 * With the plain solver the full-array problems and the subarray problems of
 * a block share one launch that takes them largest first from one list
 * (one workgroup per problem, or a persistent grid under
 * lam_set_persistent_workgroups); with a warm start the subarray problems run
 * in a second, persistent launch.  ngroups = 0 clears (and restores 2B output
 * beams). */
int lam_set_subarrays(lam_ctx *ctx, int ngroups, const uint8_t *mask, const int32_t *k);
/* Grid of the subarray eigen launches.  0 (the default): the combined
 * full-array + subarray list runs one workgroup per problem, and the
 * warm-start subarray launch is persistent with as many workgroups as fit on
 * the device at once.  > 0: both run as persistent workgroups, at most this
 * many, taking problems from an atomic queue.  Results are bit-identical
 * either way.  No counterpart in the reference; a tuning and test knob (a
 * small grid makes every workgroup take many problems from the queue). */
int lam_set_persistent_workgroups(lam_ctx *ctx, int max_workgroups);
/* Results of the last block, problem q = (c * 2 + pol) * ngroups + g, synchronous
 * host copies (any pointer may be NULL): evals [q][R] (first n_q valid,
 * ascending), evecs and proj [q][R*R][2] (n_q x n_q column-major at the start of
 * each slot), counts / sizes / sweeps [q]. */
int lam_subarray_results(lam_ctx *ctx, float *evals, float *evecs, float *proj, int32_t *counts,
                         int32_t *sizes, int32_t *sweeps);

/* Allocate and randomly initialise the MLP: warm_vectors (even, 2..8) and
 * hidden width (multiple of 32).  Called implicitly with (4, 256, 1) by the
 * other lam_learn_* calls when needed. */
int lam_learn_init(lam_ctx *ctx, int warm_vectors, int hidden, uint64_t seed);
/* dims[7] = {warm_vectors, receivers, features, padded features, hidden,
 * outputs, padded outputs}; params = FP32 master parameter count laid out as
 * W1[hidden][padded features], b1[hidden], W2[padded outputs][hidden], b2. */
int lam_learn_dims(lam_ctx *ctx, int32_t *dims, int64_t *params);
/* Record the features and the plain-Jacobi labels (top warm_vectors whitened
 * eigenvectors and the mode count K) of the next max_blocks blocks run with
 * LAM_EIG_PLAIN; lam_learn_captured returns how many blocks are held. */
int lam_learn_capture(lam_ctx *ctx, int max_blocks);
int lam_learn_captured(lam_ctx *ctx);
/* Captured training data of block `block` (0-based, < lam_learn_captured):
 * x [C*2][padded features] BF16 bits, labels [C*2][warm_vectors][receivers][2]
 * FP32 (the block's own plain-Jacobi vectors, the next block's warm-start
 * input), k [C*2].  Any pointer may be NULL.  Synchronous.  No counterpart in
 * the reference (inspection and tests of the learned warm start). */
int lam_learn_sample(lam_ctx *ctx, int block, uint16_t *x, float *labels, int32_t *k);
/* Features x [C*2][padded features] (BF16 bits) and network output y
 * [C*2][padded outputs] (FP32, skip connection included) of the most recent
 * block run with LAM_EIG_LEARNED.  Either pointer may be NULL.  Synchronous.
 * No counterpart in the reference (tests of the inference kernels). */
int lam_learn_last(lam_ctx *ctx, uint16_t *x, float *y);
/* Adam on the captured set with the project's own forward / backward
 * kernels: loss = mean over samples of sum_{k<K} (1 - |u_k^H y_k|^2/|y_k|^2).
 * loss_per_epoch (may be NULL) receives `epochs` values.  Synchronous. */
int lam_learn_train(lam_ctx *ctx, int epochs, int batch, float lr, uint64_t seed,
                    float *loss_per_epoch);
/* Read (set = 0) or replace (set = 1) the FP32 master parameters (host array
 * of the size reported by lam_learn_dims). */
int lam_learn_params(lam_ctx *ctx, float *params, int64_t count, int set);
/* Loss and gradient at the current parameters for m host samples:
 * x [m][padded features] BF16 bits, labels [m][warm_vectors][receivers][2]
 * FP32, k [m]; grad receives the parameter-shaped gradient.  Test entry. */
int lam_learn_grad(lam_ctx *ctx, const uint16_t *x, const float *labels, const int32_t *k, int m,
                   float *loss, float *grad);

/* ---- seeded synthetic voltages (benchmarks and tests) ----
 * Fills a whole device ring [C][NP+2][F][NT][RPP][2][2] and unit scales:
 * x = round(gain_c (sigma n + sum_s tone_s(t) steer[c][s][r] pol_s)), clipped
 * to +-127, with n ~ N(0,1) per real component from a counter hash of
 * (seed, block_index, sample) and t = (slot - 1) NT + sample-in-packet.
 * d_tones float[C][n_sources][6] = {amplitude, cycles/sample, phase (rad),
 * pol-0 gain, pol-1 gain re, im}; d_steer float[C][n_sources][R][2];
 * d_gain float[C].  n_sources <= 8.  Ordered on `stream`; with stream NULL the
 * call returns once the block is written. */
int lam_synth_block(const lam_shape *shape, uint64_t seed, int32_t block_index, float sigma,
                    int32_t n_sources, const float *d_tones, const float *d_steer,
                    const float *d_gain, int8_t *d_samples, int16_t *d_scales, void *stream);

/* ---- kernel-level entry points (device pointers, stream = NULL for the
 * default stream); used by unit tests the way the reference tests its
 * kernels directly (tests/test_adaptive_pipeline.cu:743-1129). ---- */

/* Batched eigen / RFI-null stage on nmat covariance matrices given as
 * float32 complex [nmat][n][n] column-major (full Hermitian), n <= 64:
 * whitening (spatial.cuh:1065-1093), Jacobi eig, ascending sort.
 * evals [nmat][n], evecs [nmat][n][n][2] column-major, sweeps [nmat] (may be NULL). */
int lam_eigh_batched(const float *d_A, int n, int nmat, int whiten, int max_sweeps,
                     float tol, float *d_evals, float *d_evecs, int32_t *d_sweeps,
                     void *stream);

/* Warm-start solve (the LAM_EIG_LEARNED / PREVIOUS refinement) of nmat
 * whitened Hermitian matrices [nmat][n][n] (column-major complex float32)
 * from proposals [nmat][warm_vectors][n]: `steps` subspace-iteration steps,
 * Rayleigh-Ritz, complement.  Writes ascending evals [nmat][n], evecs
 * [nmat][n][n][2] (columns) and ratio [nmat] = leaked interference of the k
 * largest modes over the mean remaining eigenvalue (accepted when <= tol). */
int lam_eigh_warm(const float *d_A, const float *d_prop, int n, int nmat, int warm_vectors, int k,
                  int steps, float tol, float *d_evals, float *d_evecs, float *d_ratio, void *stream);

/* Correlator of one device-resident block without a pipeline context: the
 * corr-beam pipeline's TCC launch plus trim and accumulate
 * (lambda_corr_beam_only_gpu_pipeline.hpp:314-316, src/spatial.cu:136-170).
 * d_samples / d_scales are a ring block and its scales (layouts above) for
 * `shape`; window 0 reads ring slots 0..NP-1 (spatial.cuh:254-264), window 1
 * the adaptive window from slot 1 (zero delays, spatial.cuh:1431-1436).  The
 * block's exact integer visibilities are ADDED into d_vis, int64
 * [C][R(R+1)/2][P][P][2] (raw sums, no gains), so a caller accumulates blocks
 * by calling again and clears d_vis itself.  d_flags (device int32, may be
 * NULL) is OR-ed with 1 when a scaled sample left the int8 range.  Allocates
 * its per-split partial sums (C * R(R+1)/2 * 32 B per time split) on every
 * call and returns after the stream has finished the block. */
int lam_correlate(const lam_shape *shape, int window, const void *d_samples, const void *d_scales,
                  int64_t *d_vis, int32_t *d_flags, void *stream);

/* lam_correlate without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as lam_correlate, but the per-split partial sums live in a static
 * 32 MiB device array per GPU (held until the process ends), the call only
 * enqueues on `stream` and returns without synchronising, and it calls no
 * hipMalloc / hipFree.  Fewer, longer time splits are used when the automatic
 * choice would not fit; a block needing more than 32 MiB for one split per
 * channel (C * R(R+1)/2 * 32 B: e.g. more than 504 channels at 64 receivers)
 * is rejected with LAM_ERR_UNSUPPORTED and a message on stderr.  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU (different
 * streams or threads) race; calls ordered on one stream are safe. */
int lam_correlate_noalloc(const lam_shape *shape, int window, const void *d_samples, const void *d_scales,
                          int64_t *d_vis, int32_t *d_flags, void *stream);

#ifdef __cplusplus
}
#endif
#endif
