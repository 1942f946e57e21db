// Output sink contract of the pipelines: the reference's Output interface
// (include/spatial/output.hpp:22-58) verbatim, and SingleHostMemoryOutput
// (output.hpp:60-164) with pinned buffers from lam_host_alloc.  The buffered
// writer machinery (BufferedOutput, HDF5/PSRDADA writers) is not part of this path.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <stdexcept>

#include "lam.h"
#include "spatial/packet_formats.hpp"

class Output {
 public:
  virtual ~Output() = default;
  virtual size_t register_beam_data_block(const size_t start_seq_num, const size_t end_seq_num) = 0;
  virtual size_t register_visibilities_block(const size_t start_seq_num, const size_t end_seq_num,
                                             const int num_missing_packets,
                                             const int num_total_packets) = 0;
  virtual size_t register_eigendecomposition_data_block(const size_t start_seq_num,
                                                        const size_t end_seq_num) = 0;
  virtual size_t register_fft_block(const size_t start_seq_num, const size_t end_seq_num) = 0;

  virtual void *get_beam_data_landing_pointer(const size_t block_num) = 0;
  virtual void *get_visibilities_landing_pointer(const size_t block_num) = 0;
  virtual void *get_arrivals_data_landing_pointer(const size_t block_num) = 0;
  virtual void *get_eigenvalues_data_landing_pointer(const size_t block_num) = 0;
  virtual void *get_eigenvectors_data_landing_pointer(const size_t block_num) = 0;
  virtual void *get_nulled_eigenmode_counts_landing_pointer(const size_t block_num) = 0;
  virtual void *get_beam_counts_landing_pointer(const size_t block_num) = 0;
  virtual void *get_fft_landing_pointer(const size_t block_num) = 0;

  virtual void register_beam_data_transfer_complete(const size_t block_num) = 0;
  virtual void register_visibilities_transfer_complete(const size_t block_num) = 0;
  virtual void register_arrivals_transfer_complete(const size_t block_num) = 0;
  virtual void register_eigendecomposition_data_transfer_complete(const size_t block_num) = 0;
  virtual void register_beam_counts_transfer_complete(const size_t block_num) = 0;
  virtual void register_fft_transfer_complete(const size_t block_num) = 0;
};

// One landing area per product; every register_* returns block 1 (output.hpp:86-111).
template <typename T, size_t NR_OUT_BEAMS = T::NR_BEAMS>
class SingleHostMemoryOutput : public Output {
 public:
  using BeamOutput = spatial::half_t[T::NR_CHANNELS][T::NR_POLARIZATIONS][NR_OUT_BEAMS]
                                    [T::NR_PACKETS_FOR_CORRELATION * T::NR_TIME_STEPS_PER_PACKET][2];
  using Visibilities = float[T::NR_CHANNELS][T::NR_BASELINES_UNPADDED][T::NR_POLARIZATIONS]
                            [T::NR_POLARIZATIONS][2];
  using Arrivals = bool[T::NR_CHANNELS][T::NR_PACKETS_FOR_CORRELATION + 2][T::NR_FPGA_SOURCES];
  // EigenvalueOutputType / EigenvectorOutputType (output.hpp:72-73): the
  // LambdaGPU pipeline fills [C][P][P][R] and [C][P][P][R][R]; the adaptive
  // pipeline writes its [C][P][R] and [C][P][R][R] at the start of the same
  // areas (lambda_adaptive_beamformed_spectra_pipeline.hpp:914-950).
  using Eigenvalues = typename T::EigenvalueOutputType;
  using Eigenvectors = typename T::EigenvectorOutputType;
  using EigenmodeCounts = int32_t[T::NR_CHANNELS][T::NR_POLARIZATIONS];
  // FFTOutputType (lambda_adaptive_beamformed_spectra_pipeline.hpp:77-80): per
  // output beam, 64 - 2*5 fine-channel magnitudes summed over the block.
  static constexpr size_t kFineChannels =
      T::NR_TIME_STEPS_PER_PACKET > 10 ? T::NR_TIME_STEPS_PER_PACKET - 10 : 1;
  using FFTOutput = float[T::NR_CHANNELS][T::NR_POLARIZATIONS][NR_OUT_BEAMS][kFineChannels];

  BeamOutput *beam_data;
  Visibilities *visibilities;
  Arrivals *arrivals;
  Eigenvalues *eigenvalues;
  Eigenvectors *eigenvectors;
  EigenmodeCounts *nulled_eigenmode_counts;
  FFTOutput *fft_output;
  typename T::FFTOutputType &bandpass() { return *reinterpret_cast<typename T::FFTOutputType *>(fft_output); }

  size_t register_beam_data_block(const size_t, const size_t) override { return 1; }
  size_t register_visibilities_block(const size_t, const size_t, const int missing,
                                     const int total) override {
    last_missing_packets = missing;
    last_total_packets = total;
    return 1;
  }
  size_t register_eigendecomposition_data_block(const size_t, const size_t) override { return 1; }
  size_t register_fft_block(const size_t, const size_t) override { return 1; }

  void *get_beam_data_landing_pointer(const size_t) override { return beam_data; }
  void *get_visibilities_landing_pointer(const size_t) override { return visibilities; }
  void *get_arrivals_data_landing_pointer(const size_t) override { return arrivals; }
  void *get_eigenvalues_data_landing_pointer(const size_t) override { return eigenvalues; }
  void *get_eigenvectors_data_landing_pointer(const size_t) override { return eigenvectors; }
  void *get_nulled_eigenmode_counts_landing_pointer(const size_t) override {
    return nulled_eigenmode_counts;
  }
  void *get_beam_counts_landing_pointer(const size_t) override { return nulled_eigenmode_counts; }
  void *get_fft_landing_pointer(const size_t) override { return fft_output; }

  void register_beam_data_transfer_complete(const size_t) override { ++beam_blocks_done; }
  void register_visibilities_transfer_complete(const size_t) override { ++vis_blocks_done; }
  void register_arrivals_transfer_complete(const size_t) override {}
  void register_eigendecomposition_data_transfer_complete(const size_t) override { ++eig_blocks_done; }
  void register_beam_counts_transfer_complete(const size_t) override {}
  void register_fft_transfer_complete(const size_t) override { ++fft_blocks_done; }

  int last_missing_packets = 0, last_total_packets = 0;
  // completions, counted from the HIP callback thread(s)
  std::atomic<int> beam_blocks_done{0}, vis_blocks_done{0}, eig_blocks_done{0}, fft_blocks_done{0};

  SingleHostMemoryOutput() {
    beam_data = alloc<BeamOutput>();
    visibilities = alloc<Visibilities>();
    arrivals = alloc<Arrivals>();
    eigenvalues = alloc<Eigenvalues>();
    eigenvectors = alloc<Eigenvectors>();
    nulled_eigenmode_counts = alloc<EigenmodeCounts>();
    // one landing area for either product: the adaptive spectra (FFTOutput)
    // or the LambdaGPU bandpass (T::FFTOutputType, view with bandpass())
    fft_output = static_cast<FFTOutput *>(
        lam_host_alloc(std::max(sizeof(FFTOutput), sizeof(typename T::FFTOutputType))));
    if (!fft_output) throw std::runtime_error("lam_host_alloc failed");
  }
  ~SingleHostMemoryOutput() override {
    lam_host_free(beam_data);
    lam_host_free(visibilities);
    lam_host_free(arrivals);
    lam_host_free(eigenvalues);
    lam_host_free(eigenvectors);
    lam_host_free(nulled_eigenmode_counts);
    lam_host_free(fft_output);
  }

 private:
  template <typename A>
  static A *alloc() {
    A *p = static_cast<A *>(lam_host_alloc(sizeof(A)));
    if (!p) throw std::runtime_error("lam_host_alloc failed");
    return p;
  }
};
