// Kernel-level correlator entry point: one device-resident block through the
// int8-MFMA correlator and the split accumulation, without a pipeline context.
//
// Replaces the corr-beam pipeline's TCC launch and trim + accumulate
// (lambda_corr_beam_only_gpu_pipeline.hpp:314-316, src/spatial.cu:136-170) for
// callers that hold their own device buffers and stream.  The correlator
// workgroups write one int32 partial per (channel, time split); those partials
// are the only memory besides the caller's inputs and outputs.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../include/lam.h"
#include "lam_internal.hpp"

using namespace lam;

namespace {

// Shape limits of the voltage kernels (as lam_create checks them).
int check_corr_shape(const lam_shape *sh) {
  if (!sh) return LAM_ERR_ARG;
  if (sh->nr_polarizations != 2) return LAM_ERR_UNSUPPORTED;
  if (sh->nr_channels < 1 || sh->nr_fpga_sources < 1 || sh->nr_time_steps_per_packet < 1 ||
      sh->nr_receivers_per_packet < 1 || sh->nr_packets_for_correlation < 1)
    return LAM_ERR_ARG;
  if (sh->nr_receivers != sh->nr_fpga_sources * sh->nr_receivers_per_packet) return LAM_ERR_ARG;
  if (sh->nr_receivers < 1 || sh->nr_receivers > kMaxReceivers) return LAM_ERR_UNSUPPORTED;
  if ((long)sh->nr_packets_for_correlation * sh->nr_time_steps_per_packet > 65536) return LAM_ERR_UNSUPPORTED;
  return LAM_OK;
}

// Correlator parameters of one block; vis_part and err are filled by the caller.
VoltParams corr_params(const lam_shape *sh, int window, const void *samples, const void *scales, int spw,
                       int nsplit) {
  VoltParams p{};
  p.samples = static_cast<const int8_t *>(samples);
  p.scales = static_cast<const int16_t *>(scales);
  p.C = sh->nr_channels;
  p.F = sh->nr_fpga_sources;
  p.NT = sh->nr_time_steps_per_packet;
  p.R = sh->nr_receivers;
  p.RPP = sh->nr_receivers_per_packet;
  p.NP = sh->nr_packets_for_correlation;
  p.T = p.NP * p.NT;
  p.window_base = window ? p.NT : 0;
  p.nt_log2 = -1;
  for (int l = 0; l < 30; ++l)
    if ((1 << l) == p.NT) p.nt_log2 = l;
  p.samples_per_wg = spw;
  p.nsplit = nsplit;
  p.xcd_map = (p.C % 8 == 0) ? 1 : 0;
  p.NK = (p.R + 7) / 8;
  return p;
}

// Correlate into `part` ([nsplit][C][pq][NBL][2] int32), then add the splits
// into vis (int64 [C][NBL][pq][2]).
hipError_t enqueue_correlate(const VoltParams &p, int *part, int *err, int64_t *vis, hipStream_t s) {
  VoltParams q = p;
  q.vis_part = part;
  q.err = err;
  if (hipError_t e = launch_volt(q, true, false, s)) return e;
  const long nbl = (long)p.R * (p.R + 1) / 2;
  return launch_accumulate_vis(part, p.nsplit, (long long)p.C * nbl * 8, (int)nbl,
                               reinterpret_cast<long long *>(vis), false, s);
}

}  // namespace

extern "C" int lam_correlate(const lam_shape *shape, int window, const void *d_samples, const void *d_scales,
                             int64_t *d_vis, int32_t *d_flags, void *stream) {
  if (int r = check_corr_shape(shape)) return r;
  if ((window != 0 && window != 1) || !d_samples || !d_scales || !d_vis) return LAM_ERR_ARG;
  const int T = shape->nr_packets_for_correlation * shape->nr_time_steps_per_packet;
  int spw = 0, nsplit = 0;
  pick_split(T, shape->nr_channels, 0, spw, nsplit);
  const long nbl = (long)shape->nr_receivers * (shape->nr_receivers + 1) / 2;
  const size_t part_bytes = (size_t)nsplit * shape->nr_channels * nbl * 8 * sizeof(int);
  // partial sums plus one flag word when the caller passes none
  char *buf = nullptr;
  if (hipMalloc(reinterpret_cast<void **>(&buf), part_bytes + 256) != hipSuccess) return LAM_ERR_HIP;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const VoltParams p = corr_params(shape, window, d_samples, d_scales, spw, nsplit);
  int *err = d_flags ? reinterpret_cast<int *>(d_flags) : reinterpret_cast<int *>(buf);
  hipError_t e = hipSuccess;
  if (!d_flags) e = launch_zero(buf, 256, s);
  if (e == hipSuccess) e = enqueue_correlate(p, reinterpret_cast<int *>(buf + 256), err, d_vis, s);
  // the partials are freed once the kernels that read them have finished
  const hipError_t es = hipStreamSynchronize(s);
  (void)hipFree(buf);
  return (e == hipSuccess && es == hipSuccess) ? LAM_OK : LAM_ERR_HIP;
}

// ---- allocation-free variant ----
// The per-split partial sums live in a file-scope device array, one copy per
// GPU, held until the process ends.  Its first 256 bytes take the range flag
// when the caller passes none; the partials follow.  Every call on a GPU uses
// the same array, so calls that overlap on one GPU (on different streams or
// host threads) race; calls ordered on one stream do not.
namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB per GPU
constexpr size_t kScratchFlagBytes = 256;
static __device__ __align__(256) int scratch[kScratchBytes / sizeof(int)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each device

// Device address of this device's `scratch` (looked up once per device, with
// that device current; the array's host-side name is not a device address).
int scratch_address(int dev, char **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return LAM_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return LAM_OK;
}

}  // namespace

extern "C" int lam_correlate_noalloc(const lam_shape *shape, int window, const void *d_samples,
                                     const void *d_scales, int64_t *d_vis, int32_t *d_flags, void *stream) {
  if (int r = check_corr_shape(shape)) return r;
  if ((window != 0 && window != 1) || !d_samples || !d_scales || !d_vis) return LAM_ERR_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return LAM_ERR_NO_DEVICE;
  if (dev < 0 || dev >= kMaxDevices) return LAM_ERR_UNSUPPORTED;
  char *base = nullptr;
  if (int r = scratch_address(dev, &base)) return r;
  const int C = shape->nr_channels, R = shape->nr_receivers;
  const int T = shape->nr_packets_for_correlation * shape->nr_time_steps_per_packet;
  const size_t per_split = (size_t)C * R * (R + 1) / 2 * 8 * sizeof(int);
  const size_t avail = kScratchBytes - kScratchFlagBytes;
  int spw = 0, nsplit = 0;
  pick_split(T, C, 0, spw, nsplit);
  if ((size_t)nsplit * per_split > avail) {
    // fewer, longer splits; give up when even the longest allowed split does not fit
    const size_t max_split = avail / per_split;
    if (max_split >= 1) pick_split(T, C, (int)((T + max_split - 1) / max_split), spw, nsplit);
    if (max_split < 1 || (size_t)nsplit * per_split > avail) {
      std::fprintf(stderr,
                   "lam_correlate_noalloc: %d channels x %d receivers x %d samples need %zu bytes of partial "
                   "sums, more than the %zu-byte static scratch\n",
                   C, R, T, (size_t)nsplit * per_split, avail);
      return LAM_ERR_UNSUPPORTED;
    }
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const VoltParams p = corr_params(shape, window, d_samples, d_scales, spw, nsplit);
  int *err = d_flags ? reinterpret_cast<int *>(d_flags) : reinterpret_cast<int *>(base);
  const hipError_t e = enqueue_correlate(p, reinterpret_cast<int *>(base + kScratchFlagBytes), err, d_vis, s);
  return e == hipSuccess ? LAM_OK : LAM_ERR_HIP;
}
