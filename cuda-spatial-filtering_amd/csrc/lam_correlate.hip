// Kernel-level correlator entry point: one device-resident block through the
// int8-MFMA correlator and the split accumulation, without a pipeline context.
//
// Replaces the corr-beam pipeline's TCC launch and trim + accumulate
// (lambda_corr_beam_only_gpu_pipeline.hpp:314-316, src/spatial.cu:136-170) for
// callers that hold their own device buffers and stream.  The correlator
// workgroups write one int32 partial per (channel, time split); those partials
// are the only memory besides the caller's inputs and outputs.
#include <hip/hip_runtime.h>

#include <cstring>

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
