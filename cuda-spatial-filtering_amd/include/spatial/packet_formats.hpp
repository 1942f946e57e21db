// Block types at the ProcessorState -> GPUPipeline seam.
//
// Same names, template parameters, array layouts and virtual interface as the
// reference's include/spatial/packet_formats.hpp:50-148 (FinalPacketData,
// LambdaFinalPacketData) and 265-407 (LambdaConfig); host buffers come from
// lam_host_alloc (pinned) instead of cudaHostAlloc.  Wire-format parsing
// (Ethernet/UDP headers) is not part of this path.
#pragma once
#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>

#include "lam.h"
#include "spatial/half.hpp"

struct FinalPacketData {
  uint64_t start_seq_id = 0;
  uint64_t end_seq_id = 0;
  size_t buffer_index = 0;

  virtual ~FinalPacketData() = default;
  virtual void *get_samples_ptr() = 0;
  virtual size_t get_samples_elements_size() = 0;
  virtual void *get_scales_ptr() = 0;
  virtual size_t get_scales_element_size() = 0;
  virtual bool *get_arrivals_ptr() = 0;
  virtual size_t get_arrivals_size() = 0;
  virtual void zero_missing_packets() = 0;
  virtual int get_num_missing_packets() = 0;
};

template <typename PacketSamplesType, typename PacketScalesType, size_t NR_CHANNELS,
          size_t NR_PACKETS_FOR_CORRELATION, size_t NR_RECEIVERS_PER_PACKET,
          size_t NR_POLARIZATIONS, size_t NR_FPGAS>
struct LambdaFinalPacketData : public FinalPacketData {
  using ArrivalsType = bool[NR_CHANNELS][NR_PACKETS_FOR_CORRELATION + 2][NR_FPGAS];
  PacketSamplesType *samples = nullptr;
  PacketScalesType *scales = nullptr;
  ArrivalsType *arrivals = nullptr;

  void *get_samples_ptr() override { return (void *)samples; }
  void *get_scales_ptr() override { return (void *)scales; }
  bool *get_arrivals_ptr() override { return (bool *)arrivals; }
  size_t get_samples_elements_size() override { return sizeof(PacketSamplesType); }
  size_t get_scales_element_size() override { return sizeof(PacketScalesType); }
  size_t get_arrivals_size() override { return sizeof(ArrivalsType); }

  // Missing packets are blanked through their scales, including the two guard
  // slots (packet_formats.hpp:94-112).
  void zero_missing_packets() override {
    for (size_t i = 0; i < NR_CHANNELS; ++i)
      for (int j = -1; j < static_cast<int>(NR_PACKETS_FOR_CORRELATION) + 1; ++j)
        for (size_t k = 0; k < NR_FPGAS; ++k)
          if (!arrivals[0][i][j + 1][k])
            for (size_t m = 0; m < NR_RECEIVERS_PER_PACKET; ++m)
              for (size_t n = 0; n < NR_POLARIZATIONS; ++n)
                scales[0][i][j + 1][k * NR_RECEIVERS_PER_PACKET + m][n] = 0;
  }
  // Counts only the NR_PACKETS_FOR_CORRELATION window (packet_formats.hpp:113-127).
  int get_num_missing_packets() override {
    int sum = 0;
    for (size_t i = 0; i < NR_CHANNELS; ++i)
      for (size_t j = 0; j < NR_PACKETS_FOR_CORRELATION; ++j)
        for (size_t k = 0; k < NR_FPGAS; ++k)
          if (!arrivals[0][i][j + 1][k]) ++sum;
    return sum;
  }

  LambdaFinalPacketData() {
    samples = static_cast<PacketSamplesType *>(lam_host_alloc(sizeof(PacketSamplesType)));
    scales = static_cast<PacketScalesType *>(lam_host_alloc(sizeof(PacketScalesType)));
    arrivals = static_cast<ArrivalsType *>(lam_host_alloc(sizeof(ArrivalsType)));
    if (!samples || !scales || !arrivals) throw std::runtime_error("lam_host_alloc failed");
  }
  ~LambdaFinalPacketData() override {
    lam_host_free(samples);
    lam_host_free(scales);
    lam_host_free(arrivals);
  }
  LambdaFinalPacketData(const LambdaFinalPacketData &) = delete;
  LambdaFinalPacketData &operator=(const LambdaFinalPacketData &) = delete;
};

template <size_t NR_CHANNELS_T, size_t NR_FPGA_SOURCES_T, size_t NR_TIME_STEPS_PER_PACKET_T,
          size_t NR_RECEIVERS_T, size_t NR_POLARIZATIONS_T, size_t NR_RECEIVERS_PER_PACKET_T,
          size_t NR_PACKETS_FOR_CORRELATION_T, size_t NR_BEAMS_T, size_t NR_PADDED_RECEIVERS_T,
          size_t NR_PADDED_RECEIVERS_PER_BLOCK_T, size_t NR_CORRELATED_BLOCKS_TO_ACCUMULATE_T,
          bool OVERWRITE_FPGA_ID_WITH_IP_THIRD_OCTET = false, size_t FFT_DOWNSAMPLE_FACTOR_T = 128>
struct LambdaConfig {
  static constexpr size_t NR_CHANNELS = NR_CHANNELS_T;
  static constexpr size_t NR_FPGA_SOURCES = NR_FPGA_SOURCES_T;
  static constexpr size_t NR_TIME_STEPS_PER_PACKET = NR_TIME_STEPS_PER_PACKET_T;
  static constexpr size_t NR_RECEIVERS = NR_RECEIVERS_T;
  static constexpr size_t NR_POLARIZATIONS = NR_POLARIZATIONS_T;
  static constexpr size_t NR_RECEIVERS_PER_PACKET = NR_RECEIVERS_PER_PACKET_T;
  static constexpr size_t NR_PACKETS_FOR_CORRELATION = NR_PACKETS_FOR_CORRELATION_T;
  static constexpr size_t NR_BEAMS = NR_BEAMS_T;
  static constexpr size_t NR_PADDED_RECEIVERS = NR_PADDED_RECEIVERS_T;
  static constexpr size_t NR_PADDED_RECEIVERS_PER_BLOCK = NR_PADDED_RECEIVERS_PER_BLOCK_T;
  static constexpr size_t NR_CORRELATED_BLOCKS_TO_ACCUMULATE = NR_CORRELATED_BLOCKS_TO_ACCUMULATE_T;
  static constexpr size_t NR_BASELINES = NR_PADDED_RECEIVERS * (NR_PADDED_RECEIVERS + 1) / 2;
  static constexpr size_t NR_BASELINES_UNPADDED = NR_RECEIVERS * (NR_RECEIVERS + 1) / 2;
  static constexpr size_t COMPLEX = 2;
  static constexpr size_t NR_TIME_STEPS_FOR_CORRELATION =
      NR_PACKETS_FOR_CORRELATION * NR_TIME_STEPS_PER_PACKET;
  static constexpr int FFT_DOWNSAMPLE_FACTOR = FFT_DOWNSAMPLE_FACTOR_T;
  static_assert(NR_RECEIVERS == NR_FPGA_SOURCES * NR_RECEIVERS_PER_PACKET,
                "NR_RECEIVERS must equal NR_FPGA_SOURCES * NR_RECEIVERS_PER_PACKET");

  using Sample = std::complex<int8_t>;
  using InputPacketSamplesType =
      std::complex<int8_t>[NR_CHANNELS][NR_PACKETS_FOR_CORRELATION + 2][NR_FPGA_SOURCES]
                          [NR_TIME_STEPS_PER_PACKET][NR_RECEIVERS_PER_PACKET][NR_POLARIZATIONS];
  using PacketScalesType =
      int16_t[NR_CHANNELS][NR_PACKETS_FOR_CORRELATION + 2][NR_RECEIVERS][NR_POLARIZATIONS];
  using PacketFinalDataType =
      LambdaFinalPacketData<InputPacketSamplesType, PacketScalesType, NR_CHANNELS,
                            NR_PACKETS_FOR_CORRELATION, NR_RECEIVERS_PER_PACKET, NR_POLARIZATIONS,
                            NR_FPGA_SOURCES>;
  using BeamOutputType = spatial::half_t[NR_CHANNELS][NR_POLARIZATIONS][NR_BEAMS]
                                        [NR_TIME_STEPS_FOR_CORRELATION][COMPLEX];
  using ArrivalsOutputType = bool[NR_CHANNELS][NR_PACKETS_FOR_CORRELATION + 2][NR_FPGA_SOURCES];
  // LambdaGPU bandpass: mean |FFT|^2 per run of FFT_DOWNSAMPLE_FACTOR bins
  // (packet_formats.hpp:376-379)
  using FFTOutputType = float[NR_CHANNELS][NR_POLARIZATIONS][NR_BEAMS]
                             [NR_TIME_STEPS_FOR_CORRELATION / FFT_DOWNSAMPLE_FACTOR_T];
  using VisibilitiesOutputType =
      float[NR_CHANNELS][NR_BASELINES_UNPADDED][NR_POLARIZATIONS][NR_POLARIZATIONS][COMPLEX];
  using EigenvalueOutputType = float[NR_CHANNELS][NR_POLARIZATIONS][NR_POLARIZATIONS][NR_RECEIVERS];
  using EigenvectorOutputType =
      std::complex<float>[NR_CHANNELS][NR_POLARIZATIONS][NR_POLARIZATIONS][NR_RECEIVERS][NR_RECEIVERS];
  using Complex = std::complex<float>;
  using ReceiverArray = std::array<Complex, NR_RECEIVERS>;
  using PolArray = std::array<ReceiverArray, NR_POLARIZATIONS>;
  using AntennaGains = std::array<PolArray, NR_CHANNELS>;
  using AntennaDelays = std::array<float, NR_RECEIVERS>;

  static lam_shape shape() {
    lam_shape s{};
    s.nr_channels = (int32_t)NR_CHANNELS;
    s.nr_fpga_sources = (int32_t)NR_FPGA_SOURCES;
    s.nr_time_steps_per_packet = (int32_t)NR_TIME_STEPS_PER_PACKET;
    s.nr_receivers = (int32_t)NR_RECEIVERS;
    s.nr_polarizations = (int32_t)NR_POLARIZATIONS;
    s.nr_receivers_per_packet = (int32_t)NR_RECEIVERS_PER_PACKET;
    s.nr_packets_for_correlation = (int32_t)NR_PACKETS_FOR_CORRELATION;
    s.nr_beams = (int32_t)NR_BEAMS;
    s.nr_padded_receivers = (int32_t)NR_PADDED_RECEIVERS;
    s.nr_blocks_to_accumulate = (int32_t)NR_CORRELATED_BLOCKS_TO_ACCUMULATE;
    return s;
  }
};
