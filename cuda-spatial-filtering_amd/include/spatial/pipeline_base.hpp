// GPUPipeline seam: the reference's abstract pipeline interface
// (include/spatial/pipeline_base.hpp:9-40) with the same virtuals and
// defaults, and the part of ProcessorStateBase the pipelines call
// (release_buffer, include/spatial/spatial.hpp:718-751).  Packet capture and
// reassembly are not part of this path; ProcessorStateBase here is the
// interface a reassembly engine implements to receive buffer releases.
#pragma once
#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <vector>

class Output;
struct FinalPacketData;

class ProcessorStateBase {
 public:
  virtual ~ProcessorStateBase() = default;
  // Called once per non-dummy block when the pipeline no longer reads the
  // host buffer (pipeline/common.hpp:388-396).  May run on the HIP callback thread.
  virtual void release_buffer(const size_t buffer_index) = 0;
};

class GPUPipeline {
 public:
  virtual ~GPUPipeline() = default;
  void set_state(ProcessorStateBase *state) { state_ = state; }
  void set_output(std::shared_ptr<Output> output) { output_ = output; }
  virtual void set_subpacket_delays(int *delays_subpacket) { subpacket_delays_ = delays_subpacket; }
  virtual void set_antenna_gains(std::complex<float> *gains) { gains_ = gains; }
  virtual void set_fine_delays(const float * /*delays_ns*/, double /*base_freq_hz*/,
                               double /*channel_bw_hz*/, int /*min_freq_ch*/) {}
  virtual void set_stream_permutation(const std::vector<int> & /*recv_perm*/,
                                      const std::vector<int> & /*pol_perm*/) {
    throw std::runtime_error("set_stream_permutation not supported by this pipeline type");
  }
  virtual void execute_pipeline(FinalPacketData *packet_data, const bool dummy_run = false) = 0;
  virtual void dump_visibilities(const uint64_t end_seq_num = 0) = 0;

 protected:
  ProcessorStateBase *state_ = nullptr;
  std::shared_ptr<Output> output_;
  int *subpacket_delays_ = nullptr;
  std::complex<float> *gains_ = nullptr;
};
