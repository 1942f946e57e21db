// C-ABI implementation: pipeline contexts, per-buffer streams and device
// buffers, hipGraph capture of the per-block kernel sequence, and the
// kernel-level test entry points.  Mirrors the resource layout of the
// reference pipelines (lambda_corr_beam_only_gpu_pipeline.hpp:385-592,
// lambda_adaptive_beamformed_spectra_pipeline.hpp:1014-1210) without their
// library glue: one fused voltage kernel replaces ~15 launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lam.h"
#include "lam_internal.hpp"

using namespace lam;

namespace {

struct ProfEvent {
  int kernel;
  hipEvent_t start, stop;
};

struct Buf {
  hipStream_t stream = nullptr;
  int8_t *d_samples = nullptr;  // staging for host inputs
  int16_t *d_scales = nullptr;
  __half *d_beams = nullptr;
  int *d_vis_part = nullptr;    // this block's partial visibilities [split][C][NBL][8]
  v8h *d_wop2 = nullptr;        // adaptive: this block's 2B-beam weight operands
  float *d_evals = nullptr, *d_evecs = nullptr;
  int *d_counts = nullptr, *d_sweeps = nullptr;
  hipGraphExec_t graph = nullptr;
  const void *graph_samples = nullptr, *graph_scales = nullptr;
  hipEvent_t done = nullptr;
};

}  // namespace

struct lam_ctx {
  lam_shape shape{};
  int mode = 0, nbuf = 0, device = 0;
  int C = 0, F = 0, NT = 0, R = 0, RPP = 0, NP = 0, B = 0, T = 0;
  long NBL = 0;
  int Bout = 0, NK = 0, NMT1 = 0, NMT2 = 0;
  size_t samples_bytes = 0, scales_bytes = 0, beams_bytes = 0;
  std::vector<Buf> bufs;
  int cur = 0;
  long long *d_vis_acc = nullptr;
  __half *d_weights = nullptr;
  float2 *d_gains = nullptr, *d_phase = nullptr;
  int *d_delays = nullptr, *d_recv_perm = nullptr, *d_pol_perm = nullptr;
  bool perm_set = false, delays_set = false;
  v8h *d_wop = nullptr;  // corr-beam: B-beam operands
  int *d_fixed_k = nullptr;
  int detect = 0, shrink = 0, max_sweeps = 12;
  float delta = 3.0f, tol = 1e-6f;
  int *d_err = nullptr;
  float *d_dump_f32 = nullptr;
  long long *d_dump_i64 = nullptr;
  hipEvent_t reset_done = nullptr;
  int blocks_acc = 0;
  bool use_graphs = true, graphs_dirty = true, profiling = false;
  int samples_per_wg = 0;
  std::vector<ProfEvent> prof;
  double prof_ms[LAM_K_COUNT] = {0};
  int prof_n[LAM_K_COUNT] = {0};
  std::string err;
};

namespace {

int fail(lam_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

#define HIPC(ctx, call)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(ctx, LAM_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                  \
  } while (0)

int sync_all(lam_ctx *ctx) {
  for (auto &b : ctx->bufs) HIPC(ctx, hipStreamSynchronize(b.stream));
  return LAM_OK;
}

void pick_split(lam_ctx *ctx, int &spw, int &nsplit) {
  const int T = ctx->T;
  spw = ctx->samples_per_wg;
  if (spw <= 0) {
    const int target = std::max(1, (1024 + ctx->C - 1) / ctx->C);
    spw = (T + target - 1) / target;
  }
  spw = ((spw + kChunk - 1) / kChunk) * kChunk;
  const int min_spw = (((T + kMaxSplits - 1) / kMaxSplits + kChunk - 1) / kChunk) * kChunk;
  if (spw < min_spw) spw = min_spw;
  if (spw > ((T + kChunk - 1) / kChunk) * kChunk) spw = ((T + kChunk - 1) / kChunk) * kChunk;
  nsplit = (T + spw - 1) / spw;
}

VoltParams base_volt(lam_ctx *ctx, const void *samples, const void *scales) {
  VoltParams p{};
  p.samples = static_cast<const int8_t *>(samples);
  p.scales = static_cast<const int16_t *>(scales);
  p.C = ctx->C;
  p.F = ctx->F;
  p.NT = ctx->NT;
  p.R = ctx->R;
  p.RPP = ctx->RPP;
  p.NP = ctx->NP;
  p.T = ctx->T;
  p.nt_log2 = -1;
  for (int l = 0; l < 30; ++l)
    if ((1 << l) == ctx->NT) p.nt_log2 = l;
  pick_split(ctx, p.samples_per_wg, p.nsplit);
  p.xcd_map = (ctx->C % 8 == 0) ? 1 : 0;
  p.err = ctx->d_err;
  p.NK = ctx->NK;
  return p;
}

struct Prof {
  lam_ctx *ctx;
  hipStream_t s;
  ProfEvent ev{};
  Prof(lam_ctx *c, hipStream_t st, int k) : ctx(c), s(st) {
    if (!ctx->profiling) return;
    ev.kernel = k;
    hipEventCreate(&ev.start);
    hipEventCreate(&ev.stop);
    hipEventRecord(ev.start, s);
  }
  ~Prof() {
    if (!ctx->profiling) return;
    hipEventRecord(ev.stop, s);
    ctx->prof.push_back(ev);
  }
};

// The per-block kernel sequence of one buffer (captured into a graph when enabled).
int enqueue_kernels(lam_ctx *ctx, Buf &b, const void *samples, const void *scales) {
  hipStream_t s = b.stream;
  const long long vis_count = (long long)ctx->C * ctx->NBL * 8;
  const bool atomic_acc = ctx->nbuf > 1;
  if (ctx->mode == LAM_MODE_CORR_BEAM) {
    VoltParams p = base_volt(ctx, samples, scales);
    p.window_base = 0;  // ring slots 0..NP-1 (spatial.cuh:254-264)
    p.vis_part = b.d_vis_part;
    p.wop = ctx->d_wop;
    p.nbeams_out = ctx->B;
    p.NMT = ctx->NMT1;
    p.beams = b.d_beams;
    {
      Prof pr(ctx, s, LAM_K_CORRBEAM);
      HIPC(ctx, launch_volt(p, true, true, s));
    }
    HIPC(ctx, launch_accumulate_vis(b.d_vis_part, p.nsplit, vis_count, ctx->d_vis_acc, atomic_acc, s));
    return LAM_OK;
  }
  // adaptive: correlate -> eig/null -> 2B beams
  VoltParams p = base_volt(ctx, samples, scales);
  p.window_base = ctx->NT;  // slot 1 + delay (spatial.cuh:1431-1436)
  p.delays = ctx->delays_set ? ctx->d_delays : nullptr;
  p.recv_perm = ctx->perm_set ? ctx->d_recv_perm : nullptr;
  p.pol_perm = ctx->perm_set ? ctx->d_pol_perm : nullptr;
  {
    VoltParams pc = p;
    pc.vis_part = b.d_vis_part;
    Prof pr(ctx, s, LAM_K_CORR);
    HIPC(ctx, launch_volt(pc, true, false, s));
  }
  {
    EigParams e{};
    e.C = ctx->C;
    e.R = ctx->R;
    e.B = ctx->B;
    e.vis_part = b.d_vis_part;
    e.nsplit = p.nsplit;
    e.weights = ctx->d_weights;
    e.fixed_k = ctx->d_fixed_k;
    e.detect = ctx->detect;
    e.delta = ctx->delta;
    e.shrink = ctx->shrink;
    e.max_sweeps = ctx->max_sweeps;
    e.tol = ctx->tol;
    e.NMT = ctx->NMT2;
    e.NK = ctx->NK;
    e.wop = b.d_wop2;
    e.evals = b.d_evals;
    e.evecs = b.d_evecs;
    e.counts = b.d_counts;
    e.sweeps = b.d_sweeps;
    Prof pr(ctx, s, LAM_K_EIG);
    HIPC(ctx, launch_eig(e, s));
  }
  {
    VoltParams pb = p;
    pb.wop = b.d_wop2;
    pb.nbeams_out = 2 * ctx->B;
    pb.NMT = ctx->NMT2;
    pb.beams = b.d_beams;
    Prof pr(ctx, s, LAM_K_BEAM2B);
    HIPC(ctx, launch_volt(pb, false, true, s));
  }
  HIPC(ctx, launch_accumulate_vis(b.d_vis_part, p.nsplit, vis_count, ctx->d_vis_acc, atomic_acc, s));
  return LAM_OK;
}

int run_block(lam_ctx *ctx, Buf &b, const void *samples, const void *scales) {
  const bool graphs = ctx->use_graphs && !ctx->profiling;
  if (!graphs) return enqueue_kernels(ctx, b, samples, scales);
  if (ctx->graphs_dirty) {
    for (auto &bb : ctx->bufs)
      if (bb.graph) {
        hipGraphExecDestroy(bb.graph);
        bb.graph = nullptr;
      }
    ctx->graphs_dirty = false;
  }
  if (b.graph && (b.graph_samples != samples || b.graph_scales != scales)) {
    hipGraphExecDestroy(b.graph);
    b.graph = nullptr;
  }
  if (!b.graph) {
    hipGraph_t g = nullptr;
    HIPC(ctx, hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_kernels(ctx, b, samples, scales);
    const hipError_t ec = hipStreamEndCapture(b.stream, &g);
    if (rc != LAM_OK || ec != hipSuccess || !g) {
      if (g) hipGraphDestroy(g);
      (void)hipGetLastError();
      // capture failure falls back to eager launches
      // (lambda_corr_beam_only_gpu_pipeline.hpp:346-383, 561-586)
      ctx->use_graphs = false;
      return enqueue_kernels(ctx, b, samples, scales);
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess || !ex) {
      (void)hipGetLastError();
      ctx->use_graphs = false;
      return enqueue_kernels(ctx, b, samples, scales);
    }
    b.graph = ex;
    b.graph_samples = samples;
    b.graph_scales = scales;
  }
  HIPC(ctx, hipGraphLaunch(b.graph, b.stream));
  return LAM_OK;
}

template <typename T>
int dalloc(lam_ctx *ctx, T **p, size_t bytes) {
  HIPC(ctx, hipMalloc(reinterpret_cast<void **>(p), bytes ? bytes : 16));
  HIPC(ctx, hipMemset(*p, 0, bytes ? bytes : 16));
  return LAM_OK;
}

int update_wop(lam_ctx *ctx) {
  if (ctx->mode != LAM_MODE_CORR_BEAM) return LAM_OK;
  hipStream_t s = ctx->bufs[0].stream;
  HIPC(ctx, launch_prep_weights(ctx->d_weights, ctx->d_gains, ctx->d_phase, ctx->C, ctx->B,
                                ctx->R, ctx->NMT1, ctx->NK, ctx->d_wop, s));
  HIPC(ctx, hipStreamSynchronize(s));
  return LAM_OK;
}

}  // namespace

extern "C" {

int lam_abi_version(void) { return LAM_ABI_VERSION; }

const char *lam_last_error(const lam_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lam_create(const lam_shape *sh, int mode, int num_buffers, lam_ctx **out) {
  if (!sh || !out) return LAM_ERR_ARG;
  *out = nullptr;
  if (mode != LAM_MODE_CORR_BEAM && mode != LAM_MODE_ADAPTIVE) return LAM_ERR_ARG;
  if (sh->nr_polarizations != 2) return LAM_ERR_UNSUPPORTED;
  if (sh->nr_channels < 1 || sh->nr_fpga_sources < 1 || sh->nr_time_steps_per_packet < 1 ||
      sh->nr_receivers_per_packet < 1 || sh->nr_packets_for_correlation < 1 || sh->nr_beams < 1)
    return LAM_ERR_ARG;
  if (sh->nr_receivers != sh->nr_fpga_sources * sh->nr_receivers_per_packet) return LAM_ERR_ARG;
  if (sh->nr_receivers < 1 || sh->nr_receivers > kMaxReceivers) return LAM_ERR_UNSUPPORTED;
  if (num_buffers < 1 || num_buffers > 4) return LAM_ERR_ARG;
  const long T = (long)sh->nr_packets_for_correlation * sh->nr_time_steps_per_packet;
  if (T > 65536) return LAM_ERR_UNSUPPORTED;  // int32 per-block accumulators stay exact
  const int Bout = mode == LAM_MODE_ADAPTIVE ? 2 * sh->nr_beams : sh->nr_beams;
  if (Bout > 64 || (mode == LAM_MODE_CORR_BEAM && sh->nr_beams > 32)) return LAM_ERR_UNSUPPORTED;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return LAM_ERR_NO_DEVICE;
  int dev = 0;
  hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return LAM_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return LAM_ERR_NO_DEVICE;

  lam_ctx *ctx = new lam_ctx();
  ctx->shape = *sh;
  ctx->mode = mode;
  ctx->nbuf = num_buffers;
  ctx->device = dev;
  ctx->C = sh->nr_channels;
  ctx->F = sh->nr_fpga_sources;
  ctx->NT = sh->nr_time_steps_per_packet;
  ctx->R = sh->nr_receivers;
  ctx->RPP = sh->nr_receivers_per_packet;
  ctx->NP = sh->nr_packets_for_correlation;
  ctx->B = sh->nr_beams;
  ctx->T = (int)T;
  ctx->NBL = (long)ctx->R * (ctx->R + 1) / 2;
  ctx->Bout = Bout;
  ctx->NK = (ctx->R + 7) / 8;
  ctx->NMT1 = (2 * ctx->B + 31) / 32;
  ctx->NMT2 = (4 * ctx->B + 31) / 32;
  if (ctx->NMT2 == 3) ctx->NMT2 = 4;
  ctx->samples_bytes = (size_t)ctx->C * (ctx->NP + 2) * ctx->F * ctx->NT * ctx->RPP * 4;
  ctx->scales_bytes = (size_t)ctx->C * (ctx->NP + 2) * ctx->R * 2 * 2;
  ctx->beams_bytes = (size_t)ctx->C * 2 * ctx->Bout * ctx->T * 4;
  if (const char *e = std::getenv("LAM_DISABLE_GRAPH")) ctx->use_graphs = !(e[0] == '1');

  int rc = LAM_OK;
  auto chk = [&](int r) {
    if (r != LAM_OK && rc == LAM_OK) rc = r;
    return r == LAM_OK;
  };
  const size_t vis_elems = (size_t)ctx->C * ctx->NBL * 8;
  ctx->bufs.resize(num_buffers);
  for (auto &b : ctx->bufs) {
    if (hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking) != hipSuccess) rc = LAM_ERR_HIP;
    if (hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess) rc = LAM_ERR_HIP;
    chk(dalloc(ctx, &b.d_samples, ctx->samples_bytes));
    chk(dalloc(ctx, &b.d_scales, ctx->scales_bytes));
    chk(dalloc(ctx, &b.d_beams, ctx->beams_bytes));
    chk(dalloc(ctx, &b.d_vis_part, (size_t)kMaxSplits * vis_elems * sizeof(int)));
    if (mode == LAM_MODE_ADAPTIVE) {
      chk(dalloc(ctx, &b.d_wop2, (size_t)ctx->C * 2 * ctx->NMT2 * ctx->NK * 2 * 64 * sizeof(v8h)));
      chk(dalloc(ctx, &b.d_evals, (size_t)ctx->C * 2 * ctx->R * sizeof(float)));
      chk(dalloc(ctx, &b.d_evecs, (size_t)ctx->C * 2 * ctx->R * ctx->R * 2 * sizeof(float)));
      chk(dalloc(ctx, &b.d_counts, (size_t)ctx->C * 2 * sizeof(int)));
      chk(dalloc(ctx, &b.d_sweeps, (size_t)ctx->C * 2 * sizeof(int)));
    }
  }
  chk(dalloc(ctx, &ctx->d_vis_acc, vis_elems * sizeof(long long)));
  chk(dalloc(ctx, &ctx->d_weights, (size_t)ctx->C * 2 * ctx->B * ctx->R * 2 * sizeof(__half)));
  chk(dalloc(ctx, &ctx->d_gains, (size_t)ctx->C * 2 * ctx->R * sizeof(float2)));
  chk(dalloc(ctx, &ctx->d_phase, (size_t)ctx->C * ctx->R * sizeof(float2)));
  chk(dalloc(ctx, &ctx->d_delays, (size_t)ctx->F * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_recv_perm, (size_t)ctx->R * 2 * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_pol_perm, (size_t)ctx->R * 2 * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_wop, (size_t)ctx->C * 2 * ctx->NMT1 * ctx->NK * 2 * 64 * sizeof(v8h)));
  chk(dalloc(ctx, &ctx->d_fixed_k, (size_t)ctx->C * 2 * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_err, sizeof(int)));
  chk(dalloc(ctx, &ctx->d_dump_f32, vis_elems * sizeof(float)));
  chk(dalloc(ctx, &ctx->d_dump_i64, vis_elems * sizeof(long long)));
  if (hipEventCreateWithFlags(&ctx->reset_done, hipEventDisableTiming) != hipSuccess) rc = LAM_ERR_HIP;
  if (rc == LAM_OK) {
    // unity gains and zero fine-delay phase (get_default_gains, pipeline/common.hpp:497-503)
    std::vector<float2> ones((size_t)ctx->C * 2 * ctx->R, make_float2(1.f, 0.f));
    if (hipMemcpy(ctx->d_gains, ones.data(), ones.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->d_phase, ones.data(), (size_t)ctx->C * ctx->R * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
      rc = LAM_ERR_HIP;
    hipEventRecord(ctx->reset_done, ctx->bufs[0].stream);
    if (rc == LAM_OK) rc = update_wop(ctx);
    if (rc == LAM_OK && hipDeviceSynchronize() != hipSuccess) rc = LAM_ERR_HIP;
  }
  if (rc != LAM_OK) {
    lam_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return LAM_OK;
}

int lam_destroy(lam_ctx *ctx) {
  if (!ctx) return LAM_OK;
  for (auto &b : ctx->bufs) {
    if (b.stream) hipStreamSynchronize(b.stream);
  }
  for (auto &e : ctx->prof) {
    hipEventDestroy(e.start);
    hipEventDestroy(e.stop);
  }
  for (auto &b : ctx->bufs) {
    if (b.graph) hipGraphExecDestroy(b.graph);
    if (b.done) hipEventDestroy(b.done);
    hipFree(b.d_samples);
    hipFree(b.d_scales);
    hipFree(b.d_beams);
    hipFree(b.d_vis_part);
    hipFree(b.d_wop2);
    hipFree(b.d_evals);
    hipFree(b.d_evecs);
    hipFree(b.d_counts);
    hipFree(b.d_sweeps);
    if (b.stream) hipStreamDestroy(b.stream);
  }
  hipFree(ctx->d_vis_acc);
  hipFree(ctx->d_weights);
  hipFree(ctx->d_gains);
  hipFree(ctx->d_phase);
  hipFree(ctx->d_delays);
  hipFree(ctx->d_recv_perm);
  hipFree(ctx->d_pol_perm);
  hipFree(ctx->d_wop);
  hipFree(ctx->d_fixed_k);
  hipFree(ctx->d_err);
  hipFree(ctx->d_dump_f32);
  hipFree(ctx->d_dump_i64);
  if (ctx->reset_done) hipEventDestroy(ctx->reset_done);
  delete ctx;
  return LAM_OK;
}

int lam_set_weights(lam_ctx *ctx, const uint16_t *w) {
  if (!ctx || !w) return LAM_ERR_ARG;
  if (int r = sync_all(ctx)) return r;
  HIPC(ctx, hipMemcpy(ctx->d_weights, w, (size_t)ctx->C * 2 * ctx->B * ctx->R * 2 * sizeof(uint16_t),
                      hipMemcpyHostToDevice));
  return update_wop(ctx);
}

int lam_set_gains(lam_ctx *ctx, const float *g) {
  if (!ctx || !g) return LAM_ERR_ARG;
  if (int r = sync_all(ctx)) return r;
  HIPC(ctx, hipMemcpy(ctx->d_gains, g, (size_t)ctx->C * 2 * ctx->R * sizeof(float2), hipMemcpyHostToDevice));
  return update_wop(ctx);
}

int lam_set_subpacket_delays(lam_ctx *ctx, const int32_t *d) {
  if (!ctx || !d) return LAM_ERR_ARG;
  for (int f = 0; f < ctx->F; ++f)
    if (d[f] < -ctx->NT || d[f] > ctx->NT)
      return fail(ctx, LAM_ERR_ARG, "sub-packet delay %d of FPGA %d outside [-%d, %d]", d[f], f, ctx->NT, ctx->NT);
  if (int r = sync_all(ctx)) return r;
  HIPC(ctx, hipMemcpy(ctx->d_delays, d, (size_t)ctx->F * sizeof(int), hipMemcpyHostToDevice));
  bool any = false;
  for (int f = 0; f < ctx->F; ++f) any |= d[f] != 0;
  if (any != ctx->delays_set) ctx->graphs_dirty = true;
  ctx->delays_set = any;
  return LAM_OK;
}

int lam_set_fine_delays(lam_ctx *ctx, const float *ns, double base_hz, double bw_hz, int32_t min_ch) {
  if (!ctx) return LAM_ERR_ARG;
  if (int r = sync_all(ctx)) return r;
  std::vector<float2> ph((size_t)ctx->C * ctx->R, make_float2(1.f, 0.f));
  if (ns) {
    for (int c = 0; c < ctx->C; ++c) {
      const double fc = base_hz + (double)(min_ch + c) * bw_hz;
      for (int r = 0; r < ctx->R; ++r) {
        const double phi = -2.0 * M_PI * fc * (double)ns[r] * 1e-9;
        ph[(size_t)c * ctx->R + r] = make_float2((float)std::cos(phi), (float)std::sin(phi));
      }
    }
  }
  HIPC(ctx, hipMemcpy(ctx->d_phase, ph.data(), ph.size() * sizeof(float2), hipMemcpyHostToDevice));
  return update_wop(ctx);
}

int lam_set_stream_permutation(lam_ctx *ctx, const int32_t *recv, const int32_t *pol, size_t n) {
  if (!ctx) return LAM_ERR_ARG;
  if (n != (size_t)ctx->F * ctx->RPP * 2 || !recv || !pol)
    return fail(ctx, LAM_ERR_ARG,
                "set_stream_permutation: size must equal NR_FPGA_SOURCES * NR_RECEIVERS_PER_PACKET * NR_POLARIZATIONS");
  for (size_t i = 0; i < n; ++i)
    if (recv[i] >= ctx->R || pol[i] < 0 || pol[i] > 1)
      return fail(ctx, LAM_ERR_ARG, "set_stream_permutation: entry %zu out of range", i);
  if (int r = sync_all(ctx)) return r;
  HIPC(ctx, hipMemcpy(ctx->d_recv_perm, recv, n * sizeof(int), hipMemcpyHostToDevice));
  HIPC(ctx, hipMemcpy(ctx->d_pol_perm, pol, n * sizeof(int), hipMemcpyHostToDevice));
  bool ident = true;
  for (size_t i = 0; i < n; ++i) ident &= recv[i] == (int)(i / 2) && pol[i] == (int)(i % 2);
  if (ctx->perm_set != !ident) ctx->graphs_dirty = true;
  ctx->perm_set = !ident;
  return LAM_OK;
}

int lam_set_fixed_k(lam_ctx *ctx, const int32_t *k) {
  if (!ctx || !k) return LAM_ERR_ARG;
  if (int r = sync_all(ctx)) return r;
  HIPC(ctx, hipMemcpy(ctx->d_fixed_k, k, (size_t)ctx->C * 2 * sizeof(int), hipMemcpyHostToDevice));
  return LAM_OK;
}

int lam_set_nulling(lam_ctx *ctx, int detect, float delta, int shrink) {
  if (!ctx) return LAM_ERR_ARG;
  ctx->detect = detect ? 1 : 0;
  ctx->delta = delta;
  ctx->shrink = shrink ? 1 : 0;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_set_eig_params(lam_ctx *ctx, int max_sweeps, float tol) {
  if (!ctx || max_sweeps < 0 || !(tol >= 0.0f)) return LAM_ERR_ARG;
  ctx->max_sweeps = max_sweeps;
  ctx->tol = tol;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_set_graphs(lam_ctx *ctx, int enable) {
  if (!ctx) return LAM_ERR_ARG;
  ctx->use_graphs = enable != 0;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_set_profiling(lam_ctx *ctx, int enable) {
  if (!ctx) return LAM_ERR_ARG;
  ctx->profiling = enable != 0;
  return LAM_OK;
}

int lam_set_tuning(lam_ctx *ctx, int samples_per_wg) {
  if (!ctx || samples_per_wg < 0) return LAM_ERR_ARG;
  ctx->samples_per_wg = samples_per_wg;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_kernel_times(lam_ctx *ctx, double *ms, int32_t *launches) {
  if (!ctx) return LAM_ERR_ARG;
  if (int r = sync_all(ctx)) return r;
  for (auto &e : ctx->prof) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, e.start, e.stop) == hipSuccess) {
      ctx->prof_ms[e.kernel] += t;
      ctx->prof_n[e.kernel] += 1;
    }
    hipEventDestroy(e.start);
    hipEventDestroy(e.stop);
  }
  ctx->prof.clear();
  for (int k = 0; k < LAM_K_COUNT; ++k) {
    if (ms) ms[k] = ctx->prof_ms[k];
    if (launches) launches[k] = ctx->prof_n[k];
    ctx->prof_ms[k] = 0.0;
    ctx->prof_n[k] = 0;
  }
  return LAM_OK;
}

void *lam_next_stream(lam_ctx *ctx) { return ctx ? (void *)ctx->bufs[ctx->cur].stream : nullptr; }

int lam_blocks_accumulated(const lam_ctx *ctx) { return ctx ? ctx->blocks_acc : 0; }

int lam_execute(lam_ctx *ctx, const lam_block_io *io) {
  if (!ctx || !io || !io->samples || !io->scales) return LAM_ERR_ARG;
  Buf &b = ctx->bufs[ctx->cur];
  hipStream_t s = b.stream;
  const void *smp = io->samples, *scl = io->scales;
  if (!io->inputs_on_device) {
    // H2D of the block, then release the host buffer (pipeline/common.hpp:517-528)
    HIPC(ctx, hipMemcpyAsync(b.d_samples, io->samples, ctx->samples_bytes, hipMemcpyHostToDevice, s));
    HIPC(ctx, hipMemcpyAsync(b.d_scales, io->scales, ctx->scales_bytes, hipMemcpyHostToDevice, s));
    if (io->release_fn) HIPC(ctx, hipLaunchHostFunc(s, io->release_fn, io->release_user));
    smp = b.d_samples;
    scl = b.d_scales;
  }
  // never accumulate into a visibility buffer that a dump has not reset yet
  HIPC(ctx, hipStreamWaitEvent(s, ctx->reset_done, 0));
  if (int r = run_block(ctx, b, smp, scl)) return r;
  if (io->inputs_on_device && io->release_fn)
    HIPC(ctx, hipLaunchHostFunc(s, io->release_fn, io->release_user));
  if (io->beams_dst) {
    HIPC(ctx, hipMemcpyAsync(io->beams_dst, b.d_beams, ctx->beams_bytes, hipMemcpyDefault, s));
    if (io->beams_done_fn) HIPC(ctx, hipLaunchHostFunc(s, io->beams_done_fn, io->beams_done_user));
  }
  if (ctx->mode == LAM_MODE_ADAPTIVE) {
    bool any = false;
    if (io->eigenvalues_dst) {
      HIPC(ctx, hipMemcpyAsync(io->eigenvalues_dst, b.d_evals, (size_t)ctx->C * 2 * ctx->R * sizeof(float),
                               hipMemcpyDefault, s));
      any = true;
    }
    if (io->eigenvectors_dst) {
      HIPC(ctx, hipMemcpyAsync(io->eigenvectors_dst, b.d_evecs,
                               (size_t)ctx->C * 2 * ctx->R * ctx->R * 2 * sizeof(float), hipMemcpyDefault, s));
      any = true;
    }
    if (io->counts_dst) {
      HIPC(ctx, hipMemcpyAsync(io->counts_dst, b.d_counts, (size_t)ctx->C * 2 * sizeof(int),
                               hipMemcpyDefault, s));
      any = true;
    }
    if (any && io->eigen_done_fn) HIPC(ctx, hipLaunchHostFunc(s, io->eigen_done_fn, io->eigen_done_user));
  }
  HIPC(ctx, hipEventRecord(b.done, s));
  ctx->blocks_acc += 1;
  ctx->cur = (ctx->cur + 1) % ctx->nbuf;
  return LAM_OK;
}

static int dump_common(lam_ctx *ctx, float *f32, int64_t *i64, lam_host_fn fn, void *user) {
  hipStream_t s = ctx->bufs[0].stream;
  for (auto &b : ctx->bufs) HIPC(ctx, hipStreamWaitEvent(s, b.done, 0));
  const size_t elems = (size_t)ctx->C * ctx->NBL * 8;
  HIPC(ctx, launch_dump_vis(ctx->d_vis_acc, ctx->C, ctx->R, ctx->d_gains, ctx->d_phase,
                            f32 ? ctx->d_dump_f32 : nullptr, i64 ? ctx->d_dump_i64 : nullptr, s));
  HIPC(ctx, hipEventRecord(ctx->reset_done, s));
  if (f32) HIPC(ctx, hipMemcpyAsync(f32, ctx->d_dump_f32, elems * sizeof(float), hipMemcpyDefault, s));
  if (i64) HIPC(ctx, hipMemcpyAsync(i64, ctx->d_dump_i64, elems * sizeof(long long), hipMemcpyDefault, s));
  if (fn) HIPC(ctx, hipLaunchHostFunc(s, fn, user));
  ctx->blocks_acc = 0;
  return LAM_OK;
}

int lam_dump_visibilities(lam_ctx *ctx, float *dst, lam_host_fn fn, void *user) {
  if (!ctx || !dst) return LAM_ERR_ARG;
  return dump_common(ctx, dst, nullptr, fn, user);
}

int lam_dump_visibilities_i64(lam_ctx *ctx, int64_t *dst, lam_host_fn fn, void *user) {
  if (!ctx || !dst) return LAM_ERR_ARG;
  return dump_common(ctx, nullptr, dst, fn, user);
}

int lam_sync(lam_ctx *ctx) {
  if (!ctx) return LAM_ERR_ARG;
  if (int r = sync_all(ctx)) return r;
  int flag = 0;
  HIPC(ctx, hipMemcpy(&flag, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) {
    HIPC(ctx, hipMemset(ctx->d_err, 0, sizeof(int)));
    return fail(ctx, LAM_ERR_RANGE,
                "a scaled sample left the int8 range of the exact correlator (flag %d)", flag);
  }
  return LAM_OK;
}

void *lam_device_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
  if (hipMemset(p, 0, bytes ? bytes : 16) != hipSuccess) {
    (void)hipFree(p);
    return nullptr;
  }
  return p;
}

void lam_device_free(void *p) {
  if (p) (void)hipFree(p);
}

int lam_memcpy(void *dst, const void *src, size_t bytes) {
  if (!dst || !src) return LAM_ERR_ARG;
  return hipMemcpy(dst, src, bytes, hipMemcpyDefault) == hipSuccess ? LAM_OK : LAM_ERR_HIP;
}

int lam_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  int count = 0;
  for (int d = 0; d < n; ++d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
      ++count;
  }
  return count;
}

void *lam_host_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
  std::memset(p, 0, bytes);
  return p;
}

void lam_host_free(void *p) {
  if (p) (void)hipHostFree(p);
}

int lam_eigh_batched(const float *d_A, int n, int nmat, int whiten, int max_sweeps, float tol,
                     float *d_evals, float *d_evecs, int32_t *d_sweeps, void *stream) {
  if (!d_A || n < 1 || n > kMaxReceivers || nmat < 1 || !d_evals || !d_evecs) return LAM_ERR_ARG;
  const hipError_t e = launch_eigh_batched(d_A, n, nmat, whiten, max_sweeps, tol, d_evals, d_evecs,
                                           d_sweeps, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? LAM_OK : LAM_ERR_HIP;
}

}  // extern "C"
