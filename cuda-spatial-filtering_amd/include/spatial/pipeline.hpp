// The two GPU pipelines of the reference's correlate -> beamform -> RFI-null
// path, as C++ host classes over the C ABI in lam.h:
//   LambdaCorrBeamOnlyGPUPipeline<T>
//     include/spatial/pipeline/lambda_corr_beam_only_gpu_pipeline.hpp:3-702
//   LambdaAdaptiveBeamformedSpectraPipeline<T>
//     include/spatial/pipeline/lambda_adaptive_beamformed_spectra_pipeline.hpp:3-1221
// Same constructor arguments, GPUPipeline virtuals, Output registration
// order, buffer-release contract and exception types.  All device work is one
// fused voltage kernel (corr-beam) or correlate -> eigen/null -> 2B-beam
// kernels (adaptive) inside liblam.so.
#pragma once
#include <algorithm>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "lam.h"
#include "spatial/output.hpp"
#include "spatial/packet_formats.hpp"
#include "spatial/pipeline_base.hpp"
#include "spatial/pointing.hpp"

// complex<half>[C][P][B][R] beam weights (pipeline/common.hpp:52-55).
template <typename T>
struct BeamWeightsT {
  spatial::complex_half weights[T::NR_CHANNELS][T::NR_POLARIZATIONS][T::NR_BEAMS][T::NR_RECEIVERS];
};

// Beam steering with the reference's constructor (pipeline/common.hpp:220-248).
// With no targets it is inert, as in the reference: maybe_refresh() returns
// false and the pipeline keeps its static weights (every factory of the
// reference's test harness and benchmark builds it so,
// tests/support/pipeline_harness.hpp:73-96, 145-165; apps/bench_gpu.cu:122-125).
// Target tracking needs the pointing geometry, which is not part of this
// path: a non-empty target list throws std::runtime_error.
template <typename T>
struct BeamSteering {
  using BeamWeights = BeamWeightsT<T>;

  BeamSteering() = default;
  BeamSteering(std::vector<BeamTarget> targets, std::unordered_map<int, ENUPosition> antenna_positions,
               std::unordered_map<int, int> antenna_mapping, FrequencyPlan frequency_plan, int min_freq_channel,
               ArrayLocation array_location, double update_interval_seconds, int num_buffers,
               const typename T::AntennaGains *calibration_gains = nullptr)
      : targets_(std::move(targets)),
        antenna_positions_(std::move(antenna_positions)),
        antenna_mapping_(std::move(antenna_mapping)),
        frequency_plan_(frequency_plan),
        min_freq_channel_(min_freq_channel),
        array_location_(array_location),
        update_interval_seconds_(update_interval_seconds),
        num_buffers_(num_buffers),
        calibration_gains_(calibration_gains) {
    if (!targets_.empty())
      throw std::runtime_error(
          "BeamSteering: beam target tracking (pointing geometry) is not available in this build; "
          "construct it with no targets to keep the static beam weights");
  }

  // True once real targets have been supplied (never, here).
  bool active() const { return !targets_.empty(); }
  // Returns true if new weights were enqueued: with no targets, never.
  bool maybe_refresh() { return false; }

 private:
  std::vector<BeamTarget> targets_;
  std::unordered_map<int, ENUPosition> antenna_positions_;
  std::unordered_map<int, int> antenna_mapping_;
  FrequencyPlan frequency_plan_{};
  int min_freq_channel_ = 0;
  ArrayLocation array_location_{};
  double update_interval_seconds_ = 0.0;
  int num_buffers_ = 0;
  const typename T::AntennaGains *calibration_gains_ = nullptr;
};

namespace spatial {
namespace detail {

inline void check(lam_ctx *ctx, int rc) {
  if (rc == LAM_OK) return;
  const std::string msg = ctx ? lam_last_error(ctx) : "lam call failed";
  if (rc == LAM_ERR_STATE) throw std::logic_error(msg);
  throw std::runtime_error("lam error " + std::to_string(rc) + ": " + msg);
}

struct ReleaseContext {  // BufferReleaseContext, pipeline/common.hpp:356-360
  ProcessorStateBase *state;
  size_t buffer_index;
  bool dummy_run;
};

inline void release_buffer_host_func(void *data) {
  auto *ctx = static_cast<ReleaseContext *>(data);
  if (!ctx->dummy_run && ctx->state) ctx->state->release_buffer(ctx->buffer_index);
  delete ctx;
}

struct OutputDoneContext {  // OutputTransferCompleteContext, pipeline/common.hpp:362-365
  std::shared_ptr<Output> output;
  size_t block_index;
};

inline void beam_done_host_func(void *data) {
  auto *ctx = static_cast<OutputDoneContext *>(data);
  ctx->output->register_beam_data_transfer_complete(ctx->block_index);
  delete ctx;
}

inline void fft_done_host_func(void *data) {  // pipeline/common.hpp:453-459
  auto *ctx = static_cast<OutputDoneContext *>(data);
  ctx->output->register_fft_transfer_complete(ctx->block_index);
  delete ctx;
}

inline void eig_output_done_host_func(void *data) {  // pipeline/common.hpp:437-443
  auto *ctx = static_cast<OutputDoneContext *>(data);
  ctx->output->register_eigendecomposition_data_transfer_complete(ctx->block_index);
  delete ctx;
}

inline void vis_done_host_func(void *data) {
  auto *ctx = static_cast<OutputDoneContext *>(data);
  ctx->output->register_visibilities_transfer_complete(ctx->block_index);
  delete ctx;
}

// Eigen outputs and nulled-mode counts (pipeline/common.hpp:372-390, 420-444).
struct EigenDoneContext {
  std::shared_ptr<Output> output;
  size_t eig_block, beam_block;
  bool eig_registered, beam_counts;
  void *eig_counts_dst, *beam_counts_dst;
  const int32_t *counts_src;
  size_t counts_bytes;
  std::mutex *stats_mutex;
  std::vector<int32_t> *stats_history;
};

inline void eigen_done_host_func(void *data) {
  auto *ctx = static_cast<EigenDoneContext *>(data);
  if (ctx->eig_counts_dst) std::memcpy(ctx->eig_counts_dst, ctx->counts_src, ctx->counts_bytes);
  if (ctx->beam_counts_dst) std::memcpy(ctx->beam_counts_dst, ctx->counts_src, ctx->counts_bytes);
  if (ctx->stats_mutex && ctx->stats_history) {
    std::lock_guard<std::mutex> lock(*ctx->stats_mutex);
    const size_t n = ctx->counts_bytes / sizeof(int32_t);
    ctx->stats_history->insert(ctx->stats_history->end(), ctx->counts_src, ctx->counts_src + n);
  }
  if (ctx->output) {
    if (ctx->beam_counts) ctx->output->register_beam_counts_transfer_complete(ctx->beam_block);
    if (ctx->eig_registered)
      ctx->output->register_eigendecomposition_data_transfer_complete(ctx->eig_block);
  }
  delete ctx;
}

// start_run / stop_run timing events of the benchmark ring
// (lambda_corr_beam_only_gpu_pipeline.hpp:199-201, 224, 257): recorded on the
// stream of the buffer that runs the block, around its whole enqueue.
template <size_t N>
struct RunEvents {
  hipEvent_t start[N] = {}, stop[N] = {};
  RunEvents() {
    for (size_t i = 0; i < N; ++i)
      if (hipEventCreate(&start[i]) != hipSuccess || hipEventCreate(&stop[i]) != hipSuccess)
        throw std::runtime_error("hipEventCreate failed");
  }
  ~RunEvents() {
    for (size_t i = 0; i < N; ++i) {
      if (start[i]) (void)hipEventDestroy(start[i]);
      if (stop[i]) (void)hipEventDestroy(stop[i]);
    }
  }
  RunEvents(const RunEvents &) = delete;
  RunEvents &operator=(const RunEvents &) = delete;
};

class Context {
 public:
  Context(const lam_shape &shape, int mode, int num_buffers) {
    const int rc = lam_create(&shape, mode, num_buffers, &ctx_);
    if (rc != LAM_OK)
      throw std::runtime_error("lam_create failed with status " + std::to_string(rc));
  }
  ~Context() {
    if (ctx_) lam_destroy(ctx_);
  }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  lam_ctx *get() const { return ctx_; }
  void operator()(int rc) const { check(ctx_, rc); }

 private:
  lam_ctx *ctx_ = nullptr;
};

// Calls lam_execute for one block whose io `fill` completes.  A block is
// released exactly once whatever happens: lam_execute releases it itself, also
// when it fails; if completing the io throws first (an Output callback, say),
// the release happens here before the exception leaves.
template <typename Fill>
void execute_with_release(Context &ctx, lam_block_io &io, Fill &&fill) {
  try {
    fill();
  } catch (...) {
    if (io.release_fn) io.release_fn(io.release_user);
    throw;
  }
  ctx(lam_execute(ctx.get(), &io));
}

}  // namespace detail
}  // namespace spatial

template <typename T>
class LambdaCorrBeamOnlyGPUPipeline : public GPUPipeline {
 public:
  static constexpr size_t NR_BENCHMARKING_RUNS = 100;
  using BeamWeights = BeamWeightsT<T>;

 private:
  spatial::detail::RunEvents<NR_BENCHMARKING_RUNS> run_events_;

 public:
  // the reference's benchmark ring (lambda_corr_beam_only_gpu_pipeline.hpp:199-201):
  // hipEventElapsedTime(&ms, start_run[i], stop_run[i]) is the GPU time of run i
  hipEvent_t *const start_run = run_events_.start;
  hipEvent_t *const stop_run = run_events_.stop;

  LambdaCorrBeamOnlyGPUPipeline(const int num_buffers, BeamWeightsT<T> *h_weights,
                                BeamSteering<T> beam_steering = {},
                                int nr_blocks_to_accumulate = T::NR_CORRELATED_BLOCKS_TO_ACCUMULATE)
      : num_buffers_(num_buffers),
        nr_blocks_to_accumulate_(nr_blocks_to_accumulate),
        h_weights_(h_weights),
        beam_steering_(beam_steering),
        ctx_(T::shape(), LAM_MODE_CORR_BEAM, std::min(num_buffers, 4)) {
    ctx_(lam_set_weights(ctx_.get(), reinterpret_cast<const uint16_t *>(h_weights_)));
    // warm-up run on zeroed data, then graph capture on first real use
    // (lambda_corr_beam_only_gpu_pipeline.hpp:538-586)
    typename T::PacketFinalDataType warmup_packet;
    execute_pipeline(&warmup_packet, true);
    ctx_(lam_sync(ctx_.get()));
    visibilities_start_seq_num_ = -1;
    visibilities_missing_packets_ = 0;
    num_correlation_units_integrated_ = 0;
  }

  void execute_pipeline(FinalPacketData *packet_data, const bool dummy_run = false) override {
    const size_t start_seq_num = packet_data->start_seq_id;
    const size_t end_seq_num = packet_data->end_seq_id;
    if (visibilities_start_seq_num_ == -1)
      visibilities_start_seq_num_ = static_cast<int64_t>(packet_data->start_seq_id);
    visibilities_missing_packets_ += packet_data->get_num_missing_packets();
    beam_steering_.maybe_refresh();
    if (!dummy_run && state_ == nullptr)
      throw std::logic_error("State has not been set on GPUPipeline object!");

    lam_block_io io{};
    io.samples = packet_data->get_samples_ptr();
    io.scales = packet_data->get_scales_ptr();
    io.inputs_on_device = 0;
    io.release_fn = spatial::detail::release_buffer_host_func;
    io.release_user = new spatial::detail::ReleaseContext{state_, packet_data->buffer_index, dummy_run};
    auto stream = static_cast<hipStream_t>(lam_next_stream(ctx_.get()));
    (void)hipEventRecord(start_run[benchmark_runs_done], stream);
    spatial::detail::execute_with_release(ctx_, io, [&] {
      if (output_ != nullptr && !dummy_run) {
        const size_t block_num = output_->register_beam_data_block(start_seq_num, end_seq_num);
        if (block_num != std::numeric_limits<size_t>::max()) {
          io.beams_dst = output_->get_beam_data_landing_pointer(block_num);
          io.beams_done_fn = spatial::detail::beam_done_host_func;
          io.beams_done_user = new spatial::detail::OutputDoneContext{output_, block_num};
          std::memcpy(output_->get_arrivals_data_landing_pointer(block_num),
                      packet_data->get_arrivals_ptr(), packet_data->get_arrivals_size());
          output_->register_arrivals_transfer_complete(block_num);
        }
      }
    });
    (void)hipEventRecord(stop_run[benchmark_runs_done], stream);
    if (output_ != nullptr && !dummy_run) {
      num_correlation_units_integrated_ += 1;
      if (num_correlation_units_integrated_ >= nr_blocks_to_accumulate_) dump_visibilities();
    }
    if (!dummy_run) benchmark_runs_done = (benchmark_runs_done + 1) % NR_BENCHMARKING_RUNS;
  }

  void dump_visibilities(const uint64_t end_seq_num = 0) override {
    const int total = num_correlation_units_integrated_ * static_cast<int>(
        T::NR_CHANNELS * T::NR_PACKETS_FOR_CORRELATION * T::NR_FPGA_SOURCES);
    size_t block_num = std::numeric_limits<size_t>::max();
    if (output_ != nullptr)
      block_num = output_->register_visibilities_block(
          static_cast<size_t>(visibilities_start_seq_num_), end_seq_num,
          visibilities_missing_packets_, total);
    visibilities_start_seq_num_ = -1;
    visibilities_missing_packets_ = 0;
    if (block_num != std::numeric_limits<size_t>::max()) {
      ctx_(lam_dump_visibilities(
          ctx_.get(), static_cast<float *>(output_->get_visibilities_landing_pointer(block_num)),
          spatial::detail::vis_done_host_func,
          new spatial::detail::OutputDoneContext{output_, block_num}));
    } else {
      scratch_.resize(T::NR_CHANNELS * T::NR_BASELINES_UNPADDED * T::NR_POLARIZATIONS *
                      T::NR_POLARIZATIONS * 2);
      ctx_(lam_dump_visibilities(ctx_.get(), scratch_.data(), nullptr, nullptr));
    }
    num_correlation_units_integrated_ = 0;
  }

  // Antenna gains and fine delays are stored by the reference's base class
  // only (pipeline_base.hpp:17-22): they do not change this pipeline's output.

  // Waits for all enqueued work (the tests' cudaDeviceSynchronize()).
  void synchronize() { ctx_(lam_sync(ctx_.get())); }
  lam_ctx *native_handle() const { return ctx_.get(); }
  size_t benchmark_runs_done = 0;

 private:
  int num_buffers_;
  int nr_blocks_to_accumulate_;
  BeamWeights *h_weights_;
  BeamSteering<T> beam_steering_;
  spatial::detail::Context ctx_;
  int64_t visibilities_start_seq_num_ = -1;
  int visibilities_missing_packets_ = 0;
  int num_correlation_units_integrated_ = 0;
  std::vector<float> scratch_;
};

// LambdaGPUPipeline (include/spatial/pipeline/lambda_gpu_pipeline.hpp): beams,
// accumulated visibilities and the per-block eigendecomposition of all four
// pol-pair covariances, on the delay-aligned window, with antenna gains,
// sub-packet delays, stream permutation and fine delays applied, and the
// bandpass FFT product (register_fft_block).
template <typename T>
class LambdaGPUPipeline : public GPUPipeline {
 public:
  static constexpr size_t NR_BENCHMARKING_RUNS = 100;
  using BeamWeights = BeamWeightsT<T>;

 private:
  spatial::detail::RunEvents<NR_BENCHMARKING_RUNS> run_events_;

 public:
  // the reference's benchmark ring (lambda_gpu_pipeline.hpp:444, 466, 532)
  hipEvent_t *const start_run = run_events_.start;
  hipEvent_t *const stop_run = run_events_.stop;

  LambdaGPUPipeline(const int num_buffers, BeamWeightsT<T> *h_weights, BeamSteering<T> beam_steering = {})
      : num_buffers_(std::min(num_buffers, 4)),
        h_weights_(h_weights),
        beam_steering_(beam_steering),
        ctx_(T::shape(), LAM_MODE_LAMBDA_GPU, std::min(num_buffers, 4)) {
    ctx_(lam_set_weights(ctx_.get(), reinterpret_cast<const uint16_t *>(h_weights_)));
    // bandpass FFT over the block (lambda_gpu_pipeline.hpp:814-816, 522-531)
    constexpr size_t N = T::NR_TIME_STEPS_FOR_CORRELATION;
    static_assert(N <= 8192 && (N & (N - 1)) == 0 && N % T::FFT_DOWNSAMPLE_FACTOR == 0,
                  "the LambdaGPU bandpass needs a power-of-two block length <= 8192 that "
                  "FFT_DOWNSAMPLE_FACTOR divides");
    ctx_(lam_set_bandpass(ctx_.get(), T::FFT_DOWNSAMPLE_FACTOR));
    typename T::PacketFinalDataType warmup_packet;
    execute_pipeline(&warmup_packet, true);
    ctx_(lam_sync(ctx_.get()));
    visibilities_start_seq_num_ = -1;
    visibilities_missing_packets_ = 0;
    num_correlation_units_integrated_ = 0;
  }

  void set_subpacket_delays(int *delays_subpacket) override {
    subpacket_delays_ = delays_subpacket;
    ctx_(lam_set_subpacket_delays(ctx_.get(), delays_subpacket));
  }
  void set_antenna_gains(std::complex<float> *gains) override {
    gains_ = gains;
    ctx_(lam_set_gains(ctx_.get(), reinterpret_cast<const float *>(gains)));
  }
  void set_fine_delays(const float *delays_ns, double base_freq_hz, double channel_bw_hz,
                       int min_freq_ch) override {
    ctx_(lam_set_fine_delays(ctx_.get(), delays_ns, base_freq_hz, channel_bw_hz, min_freq_ch));
  }
  void set_stream_permutation(const std::vector<int> &recv_perm, const std::vector<int> &pol_perm) override {
    constexpr size_t NR_PERM = T::NR_FPGA_SOURCES * T::NR_RECEIVERS_PER_PACKET * T::NR_POLARIZATIONS;
    if (recv_perm.size() != NR_PERM || pol_perm.size() != NR_PERM)
      throw std::runtime_error(
          "set_stream_permutation: size must equal NR_FPGA_SOURCES * "
          "NR_RECEIVERS_PER_PACKET * NR_POLARIZATIONS");
    ctx_(lam_set_stream_permutation(ctx_.get(), recv_perm.data(), pol_perm.data(), NR_PERM));
  }

  void execute_pipeline(FinalPacketData *packet_data, const bool dummy_run = false) override {
    beam_steering_.maybe_refresh();
    const size_t start_seq_num = packet_data->start_seq_id;
    const size_t end_seq_num = packet_data->end_seq_id;
    if (visibilities_start_seq_num_ == -1)
      visibilities_start_seq_num_ = static_cast<int64_t>(packet_data->start_seq_id);
    visibilities_missing_packets_ += packet_data->get_num_missing_packets();
    if (!dummy_run && state_ == nullptr)
      throw std::logic_error("State has not been set on GPUPipeline object!");
    lam_block_io io{};
    io.samples = packet_data->get_samples_ptr();
    io.scales = packet_data->get_scales_ptr();
    io.release_fn = spatial::detail::release_buffer_host_func;
    io.release_user = new spatial::detail::ReleaseContext{state_, packet_data->buffer_index, dummy_run};
    auto stream = static_cast<hipStream_t>(lam_next_stream(ctx_.get()));
    (void)hipEventRecord(start_run[benchmark_runs_done], stream);
    spatial::detail::execute_with_release(ctx_, io, [&] {
    if (output_ != nullptr && !dummy_run) {  // lambda_gpu_pipeline.hpp:534-580
      const size_t block_num = output_->register_beam_data_block(start_seq_num, end_seq_num);
      const size_t eig_block = output_->register_eigendecomposition_data_block(start_seq_num, end_seq_num);
      if (block_num != std::numeric_limits<size_t>::max()) {
        io.beams_dst = output_->get_beam_data_landing_pointer(block_num);
        io.beams_done_fn = spatial::detail::beam_done_host_func;
        io.beams_done_user = new spatial::detail::OutputDoneContext{output_, block_num};
        std::memcpy(output_->get_arrivals_data_landing_pointer(block_num),
                    packet_data->get_arrivals_ptr(), packet_data->get_arrivals_size());
        output_->register_arrivals_transfer_complete(block_num);
      }
      const size_t fft_block = output_->register_fft_block(start_seq_num, end_seq_num);
      if (fft_block != std::numeric_limits<size_t>::max()) {
        io.spectra_dst = static_cast<float *>(output_->get_fft_landing_pointer(fft_block));
        io.spectra_done_fn = spatial::detail::fft_done_host_func;
        io.spectra_done_user = new spatial::detail::OutputDoneContext{output_, fft_block};
      }
      if (eig_block != std::numeric_limits<size_t>::max()) {
        io.eigenvalues_dst = static_cast<float *>(output_->get_eigenvalues_data_landing_pointer(eig_block));
        io.eigenvectors_dst = static_cast<float *>(output_->get_eigenvectors_data_landing_pointer(eig_block));
        io.eigen_done_fn = spatial::detail::eig_output_done_host_func;
        io.eigen_done_user = new spatial::detail::OutputDoneContext{output_, eig_block};
      }
    }
    });
    (void)hipEventRecord(stop_run[benchmark_runs_done], stream);
    if (output_ != nullptr && !dummy_run) {
      num_correlation_units_integrated_ += 1;
      if (num_correlation_units_integrated_ >= static_cast<int>(T::NR_CORRELATED_BLOCKS_TO_ACCUMULATE))
        dump_visibilities();
    }
    if (!dummy_run) benchmark_runs_done = (benchmark_runs_done + 1) % NR_BENCHMARKING_RUNS;
  }

  void dump_visibilities(const uint64_t end_seq_num = 0) override {
    const int total = num_correlation_units_integrated_ * static_cast<int>(
        T::NR_CHANNELS * T::NR_PACKETS_FOR_CORRELATION * T::NR_FPGA_SOURCES);
    size_t block_num = std::numeric_limits<size_t>::max();
    if (output_ != nullptr)
      block_num = output_->register_visibilities_block(static_cast<size_t>(visibilities_start_seq_num_),
                                                       end_seq_num, visibilities_missing_packets_, total);
    visibilities_start_seq_num_ = -1;
    visibilities_missing_packets_ = 0;
    if (block_num != std::numeric_limits<size_t>::max()) {
      ctx_(lam_dump_visibilities(
          ctx_.get(), static_cast<float *>(output_->get_visibilities_landing_pointer(block_num)),
          spatial::detail::vis_done_host_func, new spatial::detail::OutputDoneContext{output_, block_num}));
    } else {
      scratch_.resize(T::NR_CHANNELS * T::NR_BASELINES_UNPADDED * T::NR_POLARIZATIONS * T::NR_POLARIZATIONS * 2);
      ctx_(lam_dump_visibilities(ctx_.get(), scratch_.data(), nullptr, nullptr));
    }
    num_correlation_units_integrated_ = 0;
  }

  void synchronize() { ctx_(lam_sync(ctx_.get())); }
  lam_ctx *native_handle() const { return ctx_.get(); }
  size_t benchmark_runs_done = 0;

 private:
  int num_buffers_;
  BeamWeights *h_weights_;
  BeamSteering<T> beam_steering_;
  spatial::detail::Context ctx_;
  int64_t visibilities_start_seq_num_ = -1;
  int visibilities_missing_packets_ = 0;
  int num_correlation_units_integrated_ = 0;
  std::vector<float> scratch_;
};

// LambdaProjectionPipeline (include/spatial/pipeline/lambda_projection_pipeline.hpp):
// per block the projector onto the NR_SIGNAL_EIGENVECTORS largest eigenvectors of
// every (channel, p, q) covariance is accumulated; every NR_RUNS_TO_AVERAGE blocks
// (and on dump_visibilities) the average is decomposed and its eigenpairs go to
// the Output's eigendecomposition landing areas.
template <typename T, int NR_SIGNAL_EIGENVECTORS, int NR_RUNS_TO_AVERAGE>
class LambdaProjectionPipeline : public GPUPipeline {
  static_assert(NR_SIGNAL_EIGENVECTORS >= 1 && NR_SIGNAL_EIGENVECTORS <= static_cast<int>(T::NR_RECEIVERS),
                "NR_SIGNAL_EIGENVECTORS must be in [1, NR_RECEIVERS]");
  static_assert(NR_RUNS_TO_AVERAGE >= 1, "NR_RUNS_TO_AVERAGE must be >= 1");

 public:
  explicit LambdaProjectionPipeline(const int num_buffers)
      : ctx_(T::shape(), LAM_MODE_PROJECTION, std::min(num_buffers, 4)) {
    ctx_(lam_set_projection(ctx_.get(), NR_SIGNAL_EIGENVECTORS, NR_RUNS_TO_AVERAGE));
    typename T::PacketFinalDataType warmup_packet;
    execute_pipeline(&warmup_packet, true);
    // the warm-up block is not part of any average
    ctx_(lam_dump_projection(ctx_.get(), nullptr, nullptr, nullptr, nullptr));
    ctx_(lam_sync(ctx_.get()));
  }

  void set_antenna_gains(std::complex<float> *gains) override {
    gains_ = gains;
    ctx_(lam_set_gains(ctx_.get(), reinterpret_cast<const float *>(gains)));
  }

  void execute_pipeline(FinalPacketData *packet_data, const bool dummy_run = false) override {
    if (!dummy_run && state_ == nullptr)
      throw std::logic_error("State has not been set on GPUPipeline object!");
    lam_block_io io{};
    io.samples = packet_data->get_samples_ptr();
    io.scales = packet_data->get_scales_ptr();
    io.release_fn = spatial::detail::release_buffer_host_func;
    io.release_user = new spatial::detail::ReleaseContext{state_, packet_data->buffer_index, dummy_run};
    spatial::detail::execute_with_release(ctx_, io, [] {});
    if (!dummy_run && lam_projection_runs(ctx_.get()) >= NR_RUNS_TO_AVERAGE)
      dump_projection(packet_data->start_seq_id, packet_data->end_seq_id);
  }

  // Flushes whatever has accumulated (lambda_projection_pipeline.hpp:452-460).
  void dump_visibilities(const uint64_t end_seq_num = 0) override {
    if (lam_projection_runs(ctx_.get()) > 0) dump_projection(0, end_seq_num);
  }

  void synchronize() { ctx_(lam_sync(ctx_.get())); }
  lam_ctx *native_handle() const { return ctx_.get(); }

 private:
  void dump_projection(const uint64_t start_seq_num, const uint64_t end_seq_num) {
    size_t block = std::numeric_limits<size_t>::max();
    if (output_ != nullptr) block = output_->register_eigendecomposition_data_block(start_seq_num, end_seq_num);
    if (block != std::numeric_limits<size_t>::max()) {
      ctx_(lam_dump_projection(ctx_.get(), static_cast<float *>(output_->get_eigenvalues_data_landing_pointer(block)),
                               static_cast<float *>(output_->get_eigenvectors_data_landing_pointer(block)),
                               spatial::detail::eig_output_done_host_func,
                               new spatial::detail::OutputDoneContext{output_, block}));
    } else {
      ctx_(lam_dump_projection(ctx_.get(), nullptr, nullptr, nullptr, nullptr));
    }
  }

  spatial::detail::Context ctx_;
};

template <typename T>
class LambdaAdaptiveBeamformedSpectraPipeline : public GPUPipeline {
 public:
  using BeamWeights = BeamWeightsT<T>;

  LambdaAdaptiveBeamformedSpectraPipeline(const int num_buffers, BeamWeightsT<T> *h_weights,
                                          const std::unordered_map<int, int> nr_signal_eigenvectors,
                                          const int min_freq_channel,
                                          BeamSteering<T> beam_steering = {},
                                          bool shrink_eigenvalues = false,
                                          bool detect_signal_eigenmodes = false,
                                          float detection_threshold_delta = 3.0f,
                                          double eigenmode_stats_interval_seconds = 10.0)
      : num_buffers_(std::min(num_buffers, 4)),
        h_weights_(h_weights),
        beam_steering_(beam_steering),
        nr_signal_eigenvectors_(nr_signal_eigenvectors),
        min_freq_channel_(min_freq_channel),
        shrink_eigenvalues_(shrink_eigenvalues),
        detect_signal_eigenmodes_(detect_signal_eigenmodes),
        detection_threshold_delta_(detection_threshold_delta),
        eigenmode_stats_interval_seconds_(eigenmode_stats_interval_seconds),
        ctx_(T::shape(), LAM_MODE_ADAPTIVE, std::min(num_buffers, 4)) {
    ctx_(lam_set_weights(ctx_.get(), reinterpret_cast<const uint16_t *>(h_weights_)));
    // fixed K per channel from the (min_freq_channel + c) map, 0 if absent (adaptive 1157-1167)
    std::vector<int32_t> k(T::NR_CHANNELS * T::NR_POLARIZATIONS, 0);
    for (size_t c = 0; c < T::NR_CHANNELS; ++c) {
      const auto it = nr_signal_eigenvectors_.find(min_freq_channel_ + static_cast<int>(c));
      const int fixed = it != nr_signal_eigenvectors_.end() ? it->second : 0;
      for (size_t p = 0; p < T::NR_POLARIZATIONS; ++p) k[c * T::NR_POLARIZATIONS + p] = fixed;
    }
    ctx_(lam_set_fixed_k(ctx_.get(), k.data()));
    ctx_(lam_set_nulling(ctx_.get(), detect_signal_eigenmodes_, detection_threshold_delta_,
                         shrink_eigenvalues_));
    // post-beam spectra: 64-point FFT per packet (adaptive 1046-1050, 808-848)
    static_assert(T::NR_TIME_STEPS_PER_PACKET == 64,
                  "the adaptive pipeline's 64-point spectra need 64 time steps per packet");
    ctx_(lam_set_spectra(ctx_.get(), 1));
    for (int i = 0; i < num_buffers_; ++i) {
      auto *p = static_cast<int32_t *>(lam_host_alloc(sizeof(int32_t) * kCounts));
      if (!p) throw std::runtime_error("lam_host_alloc failed");
      counts_staging_.push_back(p);
    }
    typename T::PacketFinalDataType warmup_packet;
    execute_pipeline(&warmup_packet, true);
    ctx_(lam_sync(ctx_.get()));
  }

  ~LambdaAdaptiveBeamformedSpectraPipeline() override {
    try {
      ctx_(lam_sync(ctx_.get()));
    } catch (...) {
    }
    for (auto *p : counts_staging_) lam_host_free(p);
  }

  void set_subpacket_delays(int *delays_subpacket) override {
    subpacket_delays_ = delays_subpacket;
    ctx_(lam_set_subpacket_delays(ctx_.get(), delays_subpacket));
  }

  void set_stream_permutation(const std::vector<int> &recv_perm,
                              const std::vector<int> &pol_perm) override {
    constexpr size_t NR_PERM = T::NR_FPGA_SOURCES * T::NR_RECEIVERS_PER_PACKET * T::NR_POLARIZATIONS;
    if (recv_perm.size() != NR_PERM || pol_perm.size() != NR_PERM)
      throw std::runtime_error(
          "set_stream_permutation: size must equal NR_FPGA_SOURCES * "
          "NR_RECEIVERS_PER_PACKET * NR_POLARIZATIONS");
    ctx_(lam_set_stream_permutation(ctx_.get(), recv_perm.data(), pol_perm.data(), NR_PERM));
  }

  void execute_pipeline(FinalPacketData *packet_data, const bool dummy_run = false) override {
    beam_steering_.maybe_refresh();
    const uint64_t start_seq_num = packet_data->start_seq_id;
    const uint64_t end_seq_num = packet_data->end_seq_id;
    if (!dummy_run && state_ == nullptr)
      throw std::logic_error("State has not been set on GPUPipeline object!");
    lam_block_io io{};
    io.samples = packet_data->get_samples_ptr();
    io.scales = packet_data->get_scales_ptr();
    io.release_fn = spatial::detail::release_buffer_host_func;
    io.release_user = new spatial::detail::ReleaseContext{state_, packet_data->buffer_index, dummy_run};
    spatial::detail::execute_with_release(ctx_, io, [&] {
    if (output_ != nullptr && !dummy_run) {
      auto *ectx = new spatial::detail::EigenDoneContext{};
      ectx->output = output_;
      ectx->counts_src = counts_staging_[current_buffer_];
      ectx->counts_bytes = sizeof(int32_t) * kCounts;
      ectx->stats_mutex = &stats_mutex_;
      ectx->stats_history = &stats_history_;
      const size_t beam_block = output_->register_beam_data_block(start_seq_num, end_seq_num);
      if (beam_block != std::numeric_limits<size_t>::max()) {
        io.beams_dst = output_->get_beam_data_landing_pointer(beam_block);
        io.beams_done_fn = spatial::detail::beam_done_host_func;
        io.beams_done_user = new spatial::detail::OutputDoneContext{output_, beam_block};
        std::memcpy(output_->get_arrivals_data_landing_pointer(beam_block),
                    packet_data->get_arrivals_ptr(), packet_data->get_arrivals_size());
        output_->register_arrivals_transfer_complete(beam_block);
        ectx->beam_block = beam_block;
        ectx->beam_counts_dst = output_->get_beam_counts_landing_pointer(beam_block);
        ectx->beam_counts = ectx->beam_counts_dst != nullptr;
      }
      const size_t eig_block = output_->register_eigendecomposition_data_block(start_seq_num, end_seq_num);
      if (eig_block != std::numeric_limits<size_t>::max()) {
        io.eigenvalues_dst = static_cast<float *>(output_->get_eigenvalues_data_landing_pointer(eig_block));
        io.eigenvectors_dst = static_cast<float *>(output_->get_eigenvectors_data_landing_pointer(eig_block));
        ectx->eig_block = eig_block;
        ectx->eig_registered = true;
        ectx->eig_counts_dst = output_->get_nulled_eigenmode_counts_landing_pointer(eig_block);
      }
      const size_t fft_block = output_->register_fft_block(start_seq_num, end_seq_num);
      if (fft_block != std::numeric_limits<size_t>::max()) {  // adaptive 900-912
        io.spectra_dst = static_cast<float *>(output_->get_fft_landing_pointer(fft_block));
        io.spectra_done_fn = spatial::detail::fft_done_host_func;
        io.spectra_done_user = new spatial::detail::OutputDoneContext{output_, fft_block};
      }
      io.counts_dst = counts_staging_[current_buffer_];
      io.eigen_done_fn = spatial::detail::eigen_done_host_func;
      io.eigen_done_user = ectx;
    }
    });
    current_buffer_ = (current_buffer_ + 1) % num_buffers_;
  }

  // The adaptive pipeline keeps no visibility product (adaptive 1209-1211).
  void dump_visibilities(const uint64_t /*end_seq_num*/ = 0) override {}

  // Per-(channel, pol) nulled-mode statistics since the last call (adaptive 959-1013).
  void print_eigenmode_stats() {
    std::vector<int32_t> samples;
    {
      std::lock_guard<std::mutex> lock(stats_mutex_);
      samples.swap(stats_history_);
    }
    if (samples.empty() || samples.size() % kCounts != 0) return;
    const size_t n = samples.size() / kCounts;
    std::cout << "Eigenmode nulling stats over last " << n << " samples" << std::endl;
    for (size_t c = 0; c < T::NR_CHANNELS; ++c)
      for (size_t p = 0; p < T::NR_POLARIZATIONS; ++p) {
        std::vector<int32_t> s;
        for (size_t i = 0; i < n; ++i) s.push_back(samples[i * kCounts + c * T::NR_POLARIZATIONS + p]);
        std::sort(s.begin(), s.end());
        double sum = 0.0;
        for (int32_t v : s) sum += v;
        const double median = s.size() % 2 == 0 ? 0.5 * (s[s.size() / 2 - 1] + s[s.size() / 2])
                                                : (double)s[s.size() / 2];
        std::cout << "  ch=" << c << " pol=" << p << " min=" << s.front() << " max=" << s.back()
                  << " median=" << median << " mean=" << sum / s.size() << std::endl;
      }
  }

  void synchronize() { ctx_(lam_sync(ctx_.get())); }
  lam_ctx *native_handle() const { return ctx_.get(); }

 private:
  static constexpr size_t kCounts = T::NR_CHANNELS * T::NR_POLARIZATIONS;
  int num_buffers_;
  BeamWeights *h_weights_;
  BeamSteering<T> beam_steering_;
  std::unordered_map<int, int> nr_signal_eigenvectors_;
  int min_freq_channel_;
  bool shrink_eigenvalues_, detect_signal_eigenmodes_;
  float detection_threshold_delta_;
  double eigenmode_stats_interval_seconds_;
  spatial::detail::Context ctx_;
  std::vector<int32_t *> counts_staging_;
  int current_buffer_ = 0;
  std::mutex stats_mutex_;
  std::vector<int32_t> stats_history_;
};
