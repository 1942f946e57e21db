// Internal declarations shared by the gfx950 kernels and the C-ABI layer.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lam {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// Voltage kernels stage 128-sample chunks through LDS.
constexpr int kChunk = 128;
constexpr int kMaxReceivers = 64;  // 4 antenna groups of 16
constexpr int kMaxSplits = 16;     // time splits per channel (partial visibility buffers)
constexpr int kEigThreads = 256;

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
  // correlator output: one int32 partial per time split [nsplit][C][NBL][2][2][2]
  int *vis_part;
  // beamformer
  const v8h *wop;          // [C][2][NMT][NK][2][64] weight operands (hi, lo)
  int nbeams_out;          // beams per (channel, pol) written
  int NMT;                 // 32-row operand tiles (16 beams each)
  int NK;                  // K steps of 8 antennas
  __half *beams;           // [C][2][nbeams_out][T][2]
  int *err;                // error flag word
};

// Parameters of the eigen / RFI-null kernel, one workgroup per (channel, pol).
struct EigParams {
  int C, R, B;
  const int *vis_part;     // per-block partial visibilities [nsplit][C][NBL][2][2][2]
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
};

// Launchers (defined in the .hip files).
hipError_t launch_volt(const VoltParams &p, bool do_corr, bool do_beam, hipStream_t s);
hipError_t launch_eig(const EigParams &p, hipStream_t s);
hipError_t launch_prep_weights(const __half *w, const float2 *gains, const float2 *phase,
                               int C, int B, int R, int NMT, int NK, v8h *wop,
                               hipStream_t s);
hipError_t launch_accumulate_vis(const int *vis_part, int nsplit, long long count, long long *acc,
                                 bool atomic, hipStream_t s);
hipError_t launch_dump_vis(long long *acc, int C, int R, const float2 *gains,
                           const float2 *phase, float *out_f32, long long *out_i64,
                           hipStream_t s);
hipError_t launch_zero(void *p, size_t bytes, hipStream_t s);
hipError_t launch_eigh_batched(const float *A, int n, int nmat, int whiten, int max_sweeps,
                               float tol, float *evals, float *evecs, int *sweeps,
                               hipStream_t s);

}  // namespace lam
