// C-ABI implementation: pipeline contexts, per-buffer streams and device
// buffers, hipGraph capture of the per-block kernel sequence, and the
// kernel-level test entry points.  Mirrors the resource layout of the
// reference pipelines (lambda_corr_beam_only_gpu_pipeline.hpp:385-592,
// lambda_adaptive_beamformed_spectra_pipeline.hpp:1014-1210) without their
// library glue: one fused voltage kernel replaces ~15 launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <numeric>
#include <random>
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
  float *d_spectra = nullptr;   // adaptive spectra [C][2][2B][NT-10] or LambdaGPU bandpass [C][2][B][T/D]
  int *d_vis_part = nullptr;    // this block's partial visibilities [split][C][pq][NBL][2]
  v8h *d_wop2 = nullptr;        // adaptive: this block's 2B-beam weight operands
  float *d_evals = nullptr, *d_evecs = nullptr;
  int *d_counts = nullptr, *d_sweeps = nullptr, *d_fallback = nullptr, *d_nsig = nullptr;
  float *d_orth = nullptr;
  hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};  // one per block phase
  int graph_full = -1;  // the eigen outputs flag captured in graph[1]
  const void *graph_samples = nullptr, *graph_scales = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t eig_done = nullptr;  // the eigen stage of this buffer's latest block
  // range flag of this buffer's blocks: cleared before each block, copied to
  // pinned host memory after it, read once the block's `done` event has passed
  int *d_err = nullptr;
  int *h_err = nullptr;
  bool err_pending = false;
  long long err_block = 0;  // index of the block whose flag is pending
};


// Learned warm start: MLP parameters, Adam state, per-block buffers and the
// captured training set.
struct Learn {
  bool ready = false;
  int KW = 4, n = 0, Dx = 0, Dp = 0, H = 256, Dy = 0, Dyp = 0;
  long P = 0, off_b1 = 0, off_W2 = 0, off_b2 = 0;
  float *d_prm = nullptr, *d_grad = nullptr, *d_m1 = nullptr, *d_m2 = nullptr;
  uint16_t *d_W1bf = nullptr, *d_W2bf = nullptr;
  int adam_t = 0;
  // per block (M = C * 2 rows)
  uint16_t *d_X = nullptr, *d_H = nullptr;
  float *d_Y = nullptr;
  float2 *d_basis = nullptr;
  // captured set
  int cap_max = 0, cap_count = 0;
  uint16_t *d_dsX = nullptr;
  float2 *d_dsU = nullptr;
  int *d_dsK = nullptr;
  // training batch scratch
  int Mb = 0;
  uint16_t *d_bX = nullptr, *d_bH = nullptr;
  float2 *d_bU = nullptr;
  int *d_bK = nullptr, *d_idx = nullptr;
  long idx_cap = 0;
  float *d_bZ = nullptr, *d_bY = nullptr, *d_bdY = nullptr, *d_bdH = nullptr, *d_loss = nullptr;
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
  void *d_spec_table = nullptr;  // DFT operand fragments of the spectra kernel
  int detect = 0, shrink = 0, max_sweeps = 12;
  float delta = 3.0f, tol = 1e-6f;
  float *d_dump_f32 = nullptr;
  long long *d_dump_i64 = nullptr;
  hipEvent_t reset_done = nullptr;
  int blocks_acc = 0;
  bool use_graphs = true, graphs_dirty = true, profiling = false;
  bool spectra = false;  // compute the post-beam spectra in the beam phase
  int bandpass = 0;      // LambdaGPU bandpass: bins averaged per output value (0: off)
  // projection mode: signal eigenvectors K, blocks per average, blocks since the last dump
  int proj_k = 1, proj_avg = 1, proj_runs = 0;
  int whiten = 1;  // adaptive: whiten the covariance before the eigen solve
  float *d_proj_acc = nullptr, *d_proj_evals = nullptr, *d_proj_evecs = nullptr;
  int samples_per_wg = 0;
  std::vector<ProfEvent> prof;
  // eigen-stage method
  int eig_method = LAM_EIG_PLAIN, refine_sweeps = 2, power = 0;
  // the current block's eigen outputs are requested (EigParams.full_outputs); set per
  // block by lam_execute, and what the autotuner times: a fresh context has none
  int eig_full = 0;
  float res_tol = 0.1f;
  int eig_plain_threads = 256;
  int wprec = kWprecF16HiLo;
  int num_cus = 256;
  // subarrays of uneven size (persistent eigen launch)
  int sub_G = 0, sub_max_n = 0;
  int *d_queue_order = nullptr, *d_queue_counter = nullptr;  // full-array + subarray problems, largest first
  int queue_n = 0;
  int persistent_wgs = 0;  // cap on the persistent launches' grid (0: as many as fit)
  uint8_t *d_sub_mask = nullptr;
  int *d_sub_k = nullptr, *d_sub_order = nullptr, *d_sub_counter = nullptr;
  float *d_sub_evals = nullptr, *d_sub_evecs = nullptr, *d_sub_proj = nullptr;
  int *d_sub_counts = nullptr, *d_sub_sweeps = nullptr, *d_sub_sizes = nullptr;
  Learn lr;
  int last_buf = -1;
  long long blocks_run = 0;     // lam_execute calls that ran a block
  long long graph_replays = 0;  // hipGraphLaunch calls
  bool graph_warned = false;
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

VoltParams base_volt(lam_ctx *ctx, const Buf &b, const void *samples, const void *scales) {
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
  // automatic split: in the adaptive mode about one correlator workgroup per CU (at
  // 384 channels one split: 0.425 ms per overlapped block against 0.438 with the
  // three splits of ~1024 workgroups, as the split count multiplies the partials the
  // eigen stage and the accumulator read); elsewhere about 1024 workgroups
  pick_split(ctx->T, ctx->C, ctx->samples_per_wg, p.samples_per_wg, p.nsplit,
             ctx->mode == LAM_MODE_ADAPTIVE ? ctx->num_cus : 1024);
  p.xcd_map = (ctx->C % 8 == 0) ? 1 : 0;
  p.err = b.d_err;
  p.NK = ctx->NK;
  p.wprec = ctx->wprec;
  return p;
}

struct Prof {
  lam_ctx *ctx;
  hipStream_t s;
  ProfEvent ev{};
  bool on = false;
  Prof(lam_ctx *c, hipStream_t st, int k) : ctx(c), s(st) {
    if (!ctx->profiling) return;
    ev.kernel = k;
    // a timing that cannot be recorded is dropped, never reported
    on = hipEventCreate(&ev.start) == hipSuccess && hipEventCreate(&ev.stop) == hipSuccess &&
         hipEventRecord(ev.start, s) == hipSuccess;
    if (!on) drop();
  }
  ~Prof() {
    if (!on) return;
    if (hipEventRecord(ev.stop, s) == hipSuccess) ctx->prof.push_back(ev);
    else drop();
  }
  void drop() {
    if (ev.start) (void)hipEventDestroy(ev.start);
    if (ev.stop) (void)hipEventDestroy(ev.stop);
    ev.start = ev.stop = nullptr;
  }
};

// Skip connection of the output layer: the previous block's vectors stored in
// the features (scaled by sqrt(n)) are added back unscaled.
Skip skip_of(const Learn &L, const uint16_t *X) {
  Skip k;
  k.X = X;
  k.ld = L.Dp;
  k.off = L.n * (L.n - 1);
  k.len = L.Dy;
  k.scale = 1.0f / std::sqrt((float)L.n);
  return k;
}

// Phases of an adaptive block: 0 = correlate, 1 = eigen stage (features, MLP,
// eigensolvers, subarrays, label capture), 2 = 2B beams + visibility
// accumulation.  Consecutive blocks on different buffers overlap; only the
// eigen stages that share state across blocks (warm-start basis, MLP scratch,
// subarray outputs) are ordered, phase 1 after the previous block's phase 1.
constexpr int kPhases = 3;

bool eig_chained(const lam_ctx *ctx) {
  return ctx->mode == LAM_MODE_ADAPTIVE && (ctx->lr.ready || ctx->sub_G > 0);
}

// Parameters of the subarray problems of one block on buffer b.
SubParams sub_params(lam_ctx *ctx, Buf &b, int nsplit) {
  SubParams q{};
  q.C = ctx->C;
  q.R = ctx->R;
  q.G = ctx->sub_G;
  q.vis_part = b.d_vis_part;
  q.nsplit = nsplit;
  q.mask = ctx->d_sub_mask;
  q.k = ctx->d_sub_k;
  q.order = ctx->d_sub_order;
  q.nprob = ctx->C * 2 * ctx->sub_G;
  q.max_n = ctx->sub_max_n;
  q.max_sweeps = ctx->max_sweeps;
  q.tol = ctx->tol;
  q.counter = ctx->d_sub_counter;
  q.evals = ctx->d_sub_evals;
  q.evecs = ctx->d_sub_evecs;
  q.proj = ctx->d_sub_proj;
  q.counts = ctx->d_sub_counts;
  q.sweeps = ctx->d_sub_sweeps;
  q.sizes = ctx->d_sub_sizes;
  // the groups' mitigated beams: rows 2B + g B + b of the 2B-beam operands
  q.weights = ctx->d_weights;
  q.B = ctx->B;
  q.NMT = ctx->NMT2;
  q.NK = ctx->NK;
  q.row0 = 2 * ctx->B;
  q.wprec = ctx->wprec;
  q.wop = b.d_wop2;
  return q;
}

// One phase of the per-block kernel sequence of one buffer (captured into a
// graph per phase when enabled).  The correlator-beamformer mode runs in phase 0.
int enqueue_phase(lam_ctx *ctx, Buf &b, const void *samples, const void *scales, int phase) {
  hipStream_t s = b.stream;
  const long long vis_count = (long long)ctx->C * ctx->NBL * 8;
  const bool atomic_acc = ctx->nbuf > 1;
  if (ctx->mode == LAM_MODE_PROJECTION) {
    if (phase != 0) return LAM_OK;
    VoltParams p = base_volt(ctx, b, samples, scales);
    p.window_base = 0;  // no alignment step: ring slots 0..NP-1 as packetToPadding reads them
    p.vis_part = b.d_vis_part;
    {
      Prof pr(ctx, s, LAM_K_CORR);
      HIPC(ctx, launch_volt(p, true, false, s));
    }
    Prof pr(ctx, s, LAM_K_POLEIG);  // lambda_projection_pipeline.hpp:380-431
    HIPC(ctx, launch_proj_block(b.d_vis_part, p.nsplit, ctx->C, ctx->R, ctx->d_gains, ctx->d_phase, ctx->proj_k,
                                ctx->max_sweeps, ctx->tol, ctx->d_proj_acc, s));
    return LAM_OK;
  }
  if (ctx->mode == LAM_MODE_CORR_BEAM || ctx->mode == LAM_MODE_LAMBDA_GPU) {
    if (phase != 0) return LAM_OK;
    VoltParams p = base_volt(ctx, b, samples, scales);
    if (ctx->mode == LAM_MODE_CORR_BEAM) {
      p.window_base = 0;  // ring slots 0..NP-1 (spatial.cuh:254-264)
    } else {
      // LambdaGPUPipeline: apply_delays + reorder_streams (lambda_gpu_pipeline.hpp:620-637),
      // the window of the adaptive pipeline: slot 1 + per-FPGA delay (spatial.cuh:1431-1436)
      p.window_base = ctx->NT;
      p.delays = ctx->delays_set ? ctx->d_delays : nullptr;
      p.recv_perm = ctx->perm_set ? ctx->d_recv_perm : nullptr;
      p.pol_perm = ctx->perm_set ? ctx->d_pol_perm : nullptr;
    }
    p.vis_part = b.d_vis_part;
    p.wop = ctx->d_wop;
    p.nbeams_out = ctx->B;
    p.NMT = ctx->NMT1;
    p.beams = b.d_beams;
    {
      Prof pr(ctx, s, LAM_K_CORRBEAM);
      HIPC(ctx, launch_volt(p, true, true, s));
    }
    if (ctx->mode == LAM_MODE_LAMBDA_GPU) {  // lambda_gpu_pipeline.hpp:494-501
      Prof pr(ctx, s, LAM_K_POLEIG);
      HIPC(ctx, launch_pol_eig(b.d_vis_part, p.nsplit, ctx->C, ctx->R, ctx->d_gains, ctx->d_phase,
                               ctx->max_sweeps, ctx->tol, b.d_evals, b.d_evecs, nullptr, s));
    }
    if (ctx->mode == LAM_MODE_LAMBDA_GPU && ctx->bandpass) {  // lambda_gpu_pipeline.hpp:522-531
      Prof pr(ctx, s, LAM_K_BANDPASS);
      HIPC(ctx, launch_bandpass(b.d_beams, ctx->C * 2 * ctx->B, ctx->T, ctx->bandpass, b.d_spectra, s));
    }
    HIPC(ctx, launch_accumulate_vis(b.d_vis_part, p.nsplit, vis_count, (int)ctx->NBL, ctx->d_vis_acc, atomic_acc, s));
    return LAM_OK;
  }
  // adaptive: correlate -> eig/null -> 2B beams
  VoltParams p = base_volt(ctx, b, samples, scales);
  p.window_base = ctx->NT;  // slot 1 + delay (spatial.cuh:1431-1436)
  p.delays = ctx->delays_set ? ctx->d_delays : nullptr;
  p.recv_perm = ctx->perm_set ? ctx->d_recv_perm : nullptr;
  p.pol_perm = ctx->perm_set ? ctx->d_pol_perm : nullptr;
  if (phase == 0) {
    VoltParams pc = p;
    pc.vis_part = b.d_vis_part;
    Prof pr(ctx, s, LAM_K_CORR);
    HIPC(ctx, launch_volt(pc, true, false, s));
    return LAM_OK;
  }
  if (phase == 2) {
    VoltParams pb = p;
    pick_split(ctx->T, ctx->C, beam_samples_per_wg(ctx->T, ctx->C, ctx->num_cus), pb.samples_per_wg, pb.nsplit);
    pb.wop = b.d_wop2;
    pb.nbeams_out = ctx->Bout;  // 2B, plus B per subarray group
    pb.NMT = ctx->NMT2;
    pb.beams = b.d_beams;
    {
      Prof pr(ctx, s, LAM_K_BEAM2B);
      HIPC(ctx, launch_volt(pb, false, true, s));
    }
    if (ctx->spectra) {  // adaptive 808-848: FFT per packet, edge bins dropped, |X| summed
      Prof pr(ctx, s, LAM_K_SPECTRA);
      HIPC(ctx, launch_spectra(b.d_beams, ctx->d_spec_table, ctx->C, ctx->Bout, ctx->T, ctx->NT, b.d_spectra, s));
    }
    HIPC(ctx, launch_accumulate_vis(b.d_vis_part, p.nsplit, vis_count, (int)ctx->NBL, ctx->d_vis_acc, atomic_acc, s));
    return LAM_OK;
  }
  const int M = ctx->C * 2;
  Learn &L = ctx->lr;
  const bool learned = ctx->eig_method == LAM_EIG_LEARNED && L.ready;
  const bool previous = ctx->eig_method == LAM_EIG_PREVIOUS && L.ready;
  const bool capture = L.ready && ctx->eig_method == LAM_EIG_PLAIN && L.cap_count < L.cap_max;
  if (learned || capture) {
    FeatParams f{};
    f.vis_part = b.d_vis_part;
    f.nsplit = p.nsplit;
    f.C = ctx->C;
    f.n = ctx->R;
    f.KW = L.KW;
    f.Dp = L.Dp;
    f.basis = L.d_basis;
    f.X = capture ? L.d_dsX + (size_t)L.cap_count * M * L.Dp : L.d_X;
    Prof pr(ctx, s, LAM_K_FEAT);
    HIPC(ctx, launch_features(f, s));
  }
  if (learned) {
    Prof pr(ctx, s, LAM_K_MLP);
    HIPC(ctx, launch_mlp_infer(L.d_X, M, L.Dp, L.d_W1bf, L.d_prm + L.off_b1, L.H, L.d_W2bf, L.d_prm + L.off_b2,
                               L.Dyp, L.d_H, L.d_Y, skip_of(L, L.d_X), s));
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
    e.whiten = ctx->whiten;
    e.full_outputs = ctx->eig_full;
    e.max_sweeps = ctx->max_sweeps;
    e.tol = ctx->tol;
    e.NMT = ctx->NMT2;
    e.NK = ctx->NK;
    e.wop = b.d_wop2;
    e.evals = b.d_evals;
    e.evecs = b.d_evecs;
    e.counts = b.d_counts;
    e.sweeps = b.d_sweeps;
    e.fallback = b.d_fallback;
    e.nsig = b.d_nsig;
    e.orth = b.d_orth;
    e.wprec = ctx->wprec;
    e.plain_threads = ctx->eig_plain_threads;
    if (L.ready) {
      e.KW = L.KW;
      e.basis_out = L.d_basis;
      e.refine_sweeps = ctx->refine_sweeps;
      e.power = ctx->power;
      e.res_tol = ctx->res_tol;
      if (learned) {
        e.proposal = reinterpret_cast<const float2 *>(L.d_Y);
        e.prop_stride = L.Dyp / 2;
      } else if (previous) {
        e.proposal = L.d_basis;
        e.prop_stride = L.KW * ctx->R;
      }
    }
    if (ctx->sub_G > 0) {
      // subarrays: every (channel, pol, group) problem in one launch; with
      // the plain solver the full-array problems join the same queue
      SubParams q = sub_params(ctx, b, p.nsplit);
      if (!eig_is_warm(e)) {
        Prof pr(ctx, s, LAM_K_EIG);
        HIPC(ctx, launch_eig_queue(e, q, ctx->d_queue_order, ctx->queue_n, ctx->d_queue_counter, ctx->num_cus,
                                   ctx->persistent_wgs, s));
      } else {
        {
          Prof pr(ctx, s, LAM_K_EIG);
          HIPC(ctx, launch_eig(e, s));
        }
        Prof pr(ctx, s, LAM_K_SUBARRAY);
        HIPC(ctx, launch_subarray_null(q, ctx->num_cus, ctx->persistent_wgs, s));
      }
    } else {
      Prof pr(ctx, s, LAM_K_EIG);
      HIPC(ctx, launch_eig(e, s));
    }
  }
  if (capture) {  // labels: this block's plain-Jacobi vectors and mode counts
    const size_t nu = (size_t)L.KW * ctx->R;
    HIPC(ctx, hipMemcpyAsync(L.d_dsU + (size_t)L.cap_count * M * nu, L.d_basis, M * nu * sizeof(float2),
                             hipMemcpyDeviceToDevice, s));
    HIPC(ctx, hipMemcpyAsync(L.d_dsK + (size_t)L.cap_count * M, b.d_nsig, M * sizeof(int),
                             hipMemcpyDeviceToDevice, s));
    L.cap_count += 1;
  }
  return LAM_OK;
}

// The whole per-block sequence on one stream (autotuner timing, no graphs).
int enqueue_kernels(lam_ctx *ctx, Buf &b, const void *samples, const void *scales) {
  for (int ph = 0; ph < kPhases; ++ph)
    if (int r = enqueue_phase(ctx, b, samples, scales, ph)) return r;
  return LAM_OK;
}

void destroy_graphs(Buf &b) {
  for (auto &g : b.graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// A failed capture or instantiation turns graphs off for the context and the
// block runs eagerly, as the reference falls back after a failed capture
// (lambda_corr_beam_only_gpu_pipeline.hpp:561-586); the warning is printed
// once per context and lam_graph_replays() stops counting.
void graph_fallback(lam_ctx *ctx, const char *what, hipError_t e) {
  (void)hipGetLastError();
  ctx->use_graphs = false;
  if (!ctx->graph_warned) {
    std::fprintf(stderr, "lam: hipGraph %s failed (%s); the pipeline runs eagerly\n", what, hipGetErrorString(e));
    ctx->graph_warned = true;
  }
}

// One phase, replayed from a graph captured on first use when graphs are on.
int run_phase(lam_ctx *ctx, Buf &b, const void *samples, const void *scales, int phase, bool graphs) {
  if (!graphs) return enqueue_phase(ctx, b, samples, scales, phase);
  hipGraphExec_t &gx = b.graph[phase];
  if (!gx) {
    hipGraph_t g = nullptr;
    HIPC(ctx, hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_phase(ctx, b, samples, scales, phase);
    const hipError_t ec = hipStreamEndCapture(b.stream, &g);
    if (rc != LAM_OK || ec != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      graph_fallback(ctx, "capture", ec != hipSuccess ? ec : hipErrorUnknown);
      return enqueue_phase(ctx, b, samples, scales, phase);
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess || !ex) {
      graph_fallback(ctx, "instantiation", ei != hipSuccess ? ei : hipErrorUnknown);
      return enqueue_phase(ctx, b, samples, scales, phase);
    }
    gx = ex;
  }
  HIPC(ctx, hipGraphLaunch(gx, b.stream));
  ctx->graph_replays += 1;
  return LAM_OK;
}

// One block: correlate; [after the previous block's eigen stage] eigen stage;
// 2B beams.  Blocks on different buffers run concurrently where they may.
int run_block(lam_ctx *ctx, Buf &b, const void *samples, const void *scales) {
  const Learn &L = ctx->lr;
  const bool capturing = L.ready && ctx->eig_method == LAM_EIG_PLAIN && L.cap_count < L.cap_max;
  const bool graphs = ctx->use_graphs && !ctx->profiling && !capturing;
  if (graphs) {
    // executable graphs are destroyed only once their last replay has finished
    if (ctx->graphs_dirty) {
      for (auto &bb : ctx->bufs) {
        HIPC(ctx, hipStreamSynchronize(bb.stream));
        destroy_graphs(bb);
      }
      ctx->graphs_dirty = false;
    }
    if (b.graph_samples != samples || b.graph_scales != scales) {
      HIPC(ctx, hipStreamSynchronize(b.stream));
      destroy_graphs(b);
      b.graph_samples = samples;
      b.graph_scales = scales;
    }
    // the eigen stage's graph holds whether eigen outputs are formed
    if (b.graph[1] && b.graph_full != ctx->eig_full) {
      HIPC(ctx, hipStreamSynchronize(b.stream));
      (void)hipGraphExecDestroy(b.graph[1]);
      b.graph[1] = nullptr;
    }
    b.graph_full = ctx->eig_full;
  }
  const bool chain = eig_chained(ctx) && ctx->last_buf >= 0 && ctx->last_buf != ctx->cur;
  for (int ph = 0; ph < kPhases; ++ph) {
    if (ph == 1 && chain) HIPC(ctx, hipStreamWaitEvent(b.stream, ctx->bufs[ctx->last_buf].eig_done, 0));
    if (int r = run_phase(ctx, b, samples, scales, ph, graphs && ctx->use_graphs)) return r;
    if (ph == 1) HIPC(ctx, hipEventRecord(b.eig_done, b.stream));
  }
  return LAM_OK;
}

template <typename T>
int dalloc(lam_ctx *ctx, T **p, size_t bytes) {
  HIPC(ctx, hipMalloc(reinterpret_cast<void **>(p), bytes ? bytes : 16));
  HIPC(ctx, hipMemset(*p, 0, bytes ? bytes : 16));
  return LAM_OK;
}

int update_wop(lam_ctx *ctx) {
  if (ctx->mode == LAM_MODE_ADAPTIVE) return LAM_OK;
  hipStream_t s = ctx->bufs[0].stream;
  HIPC(ctx, launch_prep_weights(ctx->d_weights, ctx->d_gains, ctx->d_phase, ctx->C, ctx->B,
                                ctx->R, ctx->NMT1, ctx->NK, ctx->wprec, ctx->d_wop, s));
  HIPC(ctx, hipStreamSynchronize(s));
  return LAM_OK;
}


void free_subarrays(lam_ctx *ctx) {
  for (void *q : {(void *)ctx->d_sub_mask, (void *)ctx->d_sub_k, (void *)ctx->d_sub_order,
                  (void *)ctx->d_sub_counter, (void *)ctx->d_sub_evals, (void *)ctx->d_sub_evecs,
                  (void *)ctx->d_sub_proj, (void *)ctx->d_sub_counts, (void *)ctx->d_sub_sweeps,
                  (void *)ctx->d_sub_sizes, (void *)ctx->d_queue_order, (void *)ctx->d_queue_counter})
    if (q) (void)hipFree(q);
  ctx->d_queue_order = ctx->d_queue_counter = nullptr;
  ctx->queue_n = 0;
  ctx->d_sub_mask = nullptr;
  ctx->d_sub_k = ctx->d_sub_order = ctx->d_sub_counter = nullptr;
  ctx->d_sub_evals = ctx->d_sub_evecs = ctx->d_sub_proj = nullptr;
  ctx->d_sub_counts = ctx->d_sub_sweeps = ctx->d_sub_sizes = nullptr;
  ctx->sub_G = 0;
}

// (Re)allocate every buffer's output beams, 2B-beam operands and spectra for
// ctx->Bout output beams per (channel, pol); the operands start at zero
// (rows no kernel writes stay zero).
int size_beam_outputs(lam_ctx *ctx, int Bout) {
  ctx->Bout = Bout;
  ctx->NMT2 = (2 * Bout + 31) / 32;
  if (ctx->NMT2 == 3) ctx->NMT2 = 4;
  ctx->beams_bytes = (size_t)ctx->C * 2 * ctx->Bout * ctx->T * 4;
  for (auto &b : ctx->bufs) {
    for (void *q : {(void *)b.d_beams, (void *)b.d_wop2, (void *)b.d_spectra})
      if (q) (void)hipFree(q);
    b.d_beams = nullptr;
    b.d_wop2 = nullptr;
    b.d_spectra = nullptr;
    if (int r = dalloc(ctx, &b.d_beams, ctx->beams_bytes)) return r;
    if (int r = dalloc(ctx, &b.d_wop2, (size_t)ctx->C * 2 * ctx->NMT2 * ctx->NK * 2 * 64 * sizeof(v8h))) return r;
    if (ctx->NT > 10)
      if (int r = dalloc(ctx, &b.d_spectra, (size_t)ctx->C * 2 * ctx->Bout * (ctx->NT - 10) * sizeof(float)))
        return r;
  }
  return LAM_OK;
}

void free_learn(Learn &L) {
  for (void *q : {(void *)L.d_prm, (void *)L.d_grad, (void *)L.d_m1, (void *)L.d_m2,
                  (void *)L.d_W1bf, (void *)L.d_W2bf, (void *)L.d_X, (void *)L.d_H, (void *)L.d_Y,
                  (void *)L.d_basis, (void *)L.d_dsX, (void *)L.d_dsU, (void *)L.d_dsK,
                  (void *)L.d_bX, (void *)L.d_bH, (void *)L.d_bU, (void *)L.d_bK, (void *)L.d_idx,
                  (void *)L.d_bZ, (void *)L.d_bY, (void *)L.d_bdY, (void *)L.d_bdH, (void *)L.d_loss})
    if (q) (void)hipFree(q);
  L = Learn{};
}

int round_up(int x, int m) { return (x + m - 1) / m * m; }


// Allocate the MLP and its per-block buffers; hidden-layer weights
// U(-1/sqrt(fan_in), +1/sqrt(fan_in)), zero output layer, biases and padding.
int ensure_learn(lam_ctx *ctx, int KW, int H, uint64_t seed, bool reinit) {
  Learn &L = ctx->lr;
  if (L.ready && !reinit) return LAM_OK;
  if (KW < 2 || KW > kMaxWarm || (KW & 1) || H < 32 || H % 32)
    return fail(ctx, LAM_ERR_ARG, "learn: warm_vectors must be even in [2, %d], hidden a multiple of 32", kMaxWarm);
  if (int r = sync_all(ctx)) return r;
  free_learn(L);
  const int n = ctx->R, M = ctx->C * 2;
  L.KW = KW;
  L.n = n;
  L.H = H;
  L.Dx = n * (n - 1) + 2 * KW * n;
  L.Dp = round_up(L.Dx, 32);
  L.Dy = 2 * KW * n;
  L.Dyp = round_up(L.Dy, 32);
  L.off_b1 = (long)H * L.Dp;
  L.off_W2 = L.off_b1 + H;
  L.off_b2 = L.off_W2 + (long)L.Dyp * H;
  L.P = L.off_b2 + L.Dyp;
  int rc = LAM_OK;
  auto chk = [&](int r) {
    if (r != LAM_OK && rc == LAM_OK) rc = r;
  };
  chk(dalloc(ctx, &L.d_prm, L.P * sizeof(float)));
  chk(dalloc(ctx, &L.d_grad, L.P * sizeof(float)));
  chk(dalloc(ctx, &L.d_m1, L.P * sizeof(float)));
  chk(dalloc(ctx, &L.d_m2, L.P * sizeof(float)));
  chk(dalloc(ctx, &L.d_W1bf, (size_t)H * L.Dp * sizeof(uint16_t)));
  chk(dalloc(ctx, &L.d_W2bf, (size_t)L.Dyp * H * sizeof(uint16_t)));
  chk(dalloc(ctx, &L.d_X, (size_t)M * L.Dp * sizeof(uint16_t)));
  chk(dalloc(ctx, &L.d_H, (size_t)M * H * sizeof(uint16_t)));
  chk(dalloc(ctx, &L.d_Y, (size_t)M * L.Dyp * sizeof(float)));
  chk(dalloc(ctx, &L.d_basis, (size_t)M * KW * n * sizeof(float2)));
  chk(dalloc(ctx, &L.d_loss, sizeof(float)));
  if (rc != LAM_OK) {
    free_learn(L);
    return rc;
  }
  std::vector<float> prm(L.P, 0.0f);
  std::mt19937_64 gen(seed);
  std::uniform_real_distribution<float> u1(-1.0f / std::sqrt((float)L.Dx), 1.0f / std::sqrt((float)L.Dx));
  for (int h = 0; h < H; ++h)
    for (int d = 0; d < L.Dx; ++d) prm[(size_t)h * L.Dp + d] = u1(gen);
  // the output layer starts at zero: the untrained network proposes the
  // previous block's vectors through the skip connection
  HIPC(ctx, hipMemcpy(L.d_prm, prm.data(), L.P * sizeof(float), hipMemcpyHostToDevice));
  hipStream_t st = ctx->bufs[0].stream;
  HIPC(ctx, launch_to_bf16(L.d_prm, (long)H * L.Dp, L.d_W1bf, st));
  HIPC(ctx, launch_to_bf16(L.d_prm + L.off_W2, (long)L.Dyp * H, L.d_W2bf, st));
  HIPC(ctx, hipStreamSynchronize(st));
  L.ready = true;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int ensure_batch(lam_ctx *ctx, int Mb) {
  Learn &L = ctx->lr;
  if (L.Mb >= Mb) return LAM_OK;
  for (void *q : {(void *)L.d_bX, (void *)L.d_bH, (void *)L.d_bU, (void *)L.d_bK, (void *)L.d_bZ,
                  (void *)L.d_bY, (void *)L.d_bdY, (void *)L.d_bdH})
    if (q) (void)hipFree(q);
  L.d_bX = nullptr;
  L.d_bH = nullptr;
  L.d_bU = nullptr;
  L.d_bK = nullptr;
  L.d_bZ = L.d_bY = L.d_bdY = L.d_bdH = nullptr;
  L.Mb = 0;
  if (int r = dalloc(ctx, &L.d_bX, (size_t)Mb * L.Dp * sizeof(uint16_t))) return r;
  if (int r = dalloc(ctx, &L.d_bH, (size_t)Mb * L.H * sizeof(uint16_t))) return r;
  if (int r = dalloc(ctx, &L.d_bU, (size_t)Mb * L.KW * L.n * sizeof(float2))) return r;
  if (int r = dalloc(ctx, &L.d_bK, (size_t)Mb * sizeof(int))) return r;
  if (int r = dalloc(ctx, &L.d_bZ, (size_t)Mb * L.H * sizeof(float))) return r;
  if (int r = dalloc(ctx, &L.d_bdH, (size_t)Mb * L.H * sizeof(float))) return r;
  if (int r = dalloc(ctx, &L.d_bY, (size_t)Mb * L.Dyp * sizeof(float))) return r;
  if (int r = dalloc(ctx, &L.d_bdY, (size_t)Mb * L.Dyp * sizeof(float))) return r;
  L.Mb = Mb;
  return LAM_OK;
}

// Forward, loss and backward of one batch held in the batch buffers; fills d_grad.
int forward_backward(lam_ctx *ctx, int nb, hipStream_t s) {
  Learn &L = ctx->lr;
  HIPC(ctx, launch_dense_fwd(L.d_bX, nb, L.Dp, L.d_W1bf, L.d_prm + L.off_b1, L.H, L.d_bZ, L.d_bH, Skip{}, s));
  HIPC(ctx, launch_dense_fwd(L.d_bH, nb, L.H, L.d_W2bf, L.d_prm + L.off_b2, L.Dyp, L.d_bY, nullptr,
                             skip_of(L, L.d_bX), s));
  HIPC(ctx, launch_loss(L.d_bY, L.Dyp, L.d_bU, L.d_bK, nb, L.n, L.KW, L.d_bdY, L.d_loss, s));
  // dW2[o][h] = sum_m dY[m][o] H[m][h];  db2 = colsum(dY)
  HIPC(ctx, launch_gemm_bwd(true, L.d_bdY, L.Dyp, L.d_bH, L.H, L.d_grad + L.off_W2, L.H, L.Dyp, L.H, nb,
                            nullptr, 0, s));
  HIPC(ctx, launch_colsum(L.d_bdY, nb, L.Dyp, L.d_grad + L.off_b2, s));
  // dH[m][h] = sum_o dY[m][o] W2[o][h] * (Z1[m][h] > 0)
  HIPC(ctx, launch_gemm_bwd(false, L.d_bdY, L.Dyp, L.d_W2bf, L.H, L.d_bdH, L.H, nb, L.H, L.Dyp, L.d_bZ,
                            L.H, s));
  // dW1[h][d] = sum_m dH[m][h] X[m][d];  db1 = colsum(dH)
  HIPC(ctx, launch_gemm_bwd(true, L.d_bdH, L.H, L.d_bX, L.Dp, L.d_grad, L.Dp, L.H, L.Dp, nb, nullptr, 0, s));
  HIPC(ctx, launch_colsum(L.d_bdH, nb, L.H, L.d_grad + L.off_b1, s));
  return LAM_OK;
}


// Drain the profiling events into per-kernel sums (ms) and launch counts.
int drain_prof(lam_ctx *ctx, double *ms, int *n) {
  if (int r = sync_all(ctx)) return r;
  for (auto &e : ctx->prof) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, e.start, e.stop) == hipSuccess) {
      ms[e.kernel] += t;
      n[e.kernel] += 1;
    }
    (void)hipEventDestroy(e.start);
    (void)hipEventDestroy(e.stop);
  }
  ctx->prof.clear();
  return LAM_OK;
}

// Average milliseconds per block of the kernels in `ids` over `iters` passes of
// the block sequence (after one untimed pass).
int time_sequence(lam_ctx *ctx, const void *const *smp, const void *const *scl, int nblocks, int iters,
                  std::initializer_list<int> ids, double &out) {
  double ms[LAM_K_COUNT] = {0};
  int n[LAM_K_COUNT] = {0};
  Buf &b = ctx->bufs[0];
  for (int pass = 0; pass <= iters; ++pass) {
    for (int i = 0; i < nblocks; ++i)
      if (int r = enqueue_kernels(ctx, b, smp[i], scl[i])) return r;
    if (pass == 0) {  // discard the untimed pass
      double d0[LAM_K_COUNT] = {0};
      int d1[LAM_K_COUNT] = {0};
      if (int r = drain_prof(ctx, d0, d1)) return r;
    }
  }
  if (int r = drain_prof(ctx, ms, n)) return r;
  double t = 0.0;
  for (int k : ids) t += ms[k];
  out = t / ((double)iters * nblocks);
  return LAM_OK;
}

// Streaming milliseconds per block: the blocks run back to back through all
// buffers as lam_execute runs them (eager launches, consecutive blocks
// overlapping), `iters` passes after one untimed pass.
int time_stream(lam_ctx *ctx, const void *const *smp, const void *const *scl, int nblocks, int iters,
                double &out) {
  const bool prof = ctx->profiling, graphs = ctx->use_graphs;
  ctx->profiling = false;
  ctx->use_graphs = false;
  int rc = LAM_OK;
  auto t0 = std::chrono::steady_clock::now();
  for (int pass = 0; pass <= iters && rc == LAM_OK; ++pass) {
    if (pass == 1) {
      if ((rc = sync_all(ctx))) break;
      t0 = std::chrono::steady_clock::now();
    }
    for (int i = 0; i < nblocks && rc == LAM_OK; ++i) {
      Buf &b = ctx->bufs[ctx->cur];
      if ((rc = run_block(ctx, b, smp[i], scl[i]))) break;
      if (hipEventRecord(b.done, b.stream) != hipSuccess) rc = fail(ctx, LAM_ERR_HIP, "autotune: event record");
      ctx->last_buf = ctx->cur;
      ctx->cur = (ctx->cur + 1) % ctx->nbuf;
    }
  }
  if (rc == LAM_OK) rc = sync_all(ctx);
  const auto t1 = std::chrono::steady_clock::now();
  out = std::chrono::duration<double, std::milli>(t1 - t0).count() / ((double)iters * nblocks);
  ctx->profiling = prof;
  ctx->use_graphs = graphs;
  return rc;
}

// Choice among `cands` by time per block.  The candidates are measured in
// `rounds` interleaved rounds (the order rotated every round) so that slow
// drifts of the GPU clock favour none of them; each is judged by its median.
// The incumbent (the setting in force before tuning) is kept unless the best
// median beats it by more than the incumbent's own spread over the rounds.
int pick_by_time(const std::vector<int> &cands, int incumbent, int rounds, const std::function<void(int)> &apply,
                 const std::function<int(double &)> &measure, int &chosen) {
  const size_t n = cands.size();
  std::vector<std::vector<double>> t(n);
  for (int r = 0; r < rounds; ++r)
    for (size_t j = 0; j < n; ++j) {
      const size_t i = (j + (size_t)r) % n;
      apply(cands[i]);
      double v = 0.0;
      if (int rc = measure(v)) return rc;
      t[i].push_back(v);
    }
  auto median = [](std::vector<double> v) {
    std::sort(v.begin(), v.end());
    const size_t m = v.size() / 2;
    return v.size() % 2 ? v[m] : 0.5 * (v[m - 1] + v[m]);
  };
  size_t best = 0, inc = n;
  for (size_t i = 0; i < n; ++i) {
    if (median(t[i]) < median(t[best])) best = i;
    if (cands[i] == incumbent) inc = i;
  }
  chosen = cands[best];
  if (inc < n && inc != best) {
    const double spread = *std::max_element(t[inc].begin(), t[inc].end()) -
                          *std::min_element(t[inc].begin(), t[inc].end());
    if (median(t[best]) >= median(t[inc]) - spread) chosen = incumbent;
  }
  apply(chosen);
  return LAM_OK;
}

// Reports a range error of a completed block: each block's flag is copied to
// pinned host memory at its end; the flag is read once the block's `done`
// event has passed (wait = true: after synchronising every stream).
int check_range(lam_ctx *ctx, bool wait) {
  long long bad = -1;
  for (auto &b : ctx->bufs) {
    if (!b.err_pending) continue;
    if (!wait) {
      const hipError_t q = hipEventQuery(b.done);
      if (q == hipErrorNotReady) {
        (void)hipGetLastError();
        continue;
      }
      if (q != hipSuccess) return fail(ctx, LAM_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(q));
    }
    b.err_pending = false;
    if (*b.h_err) {
      *b.h_err = 0;
      if (bad < 0 || b.err_block < bad) bad = b.err_block;
    }
  }
  if (bad >= 0)
    return fail(ctx, LAM_ERR_RANGE,
                "block %lld: a scaled sample left the int8 range of the exact correlator; its visibilities "
                "and beams are wrong",
                bad);
  return LAM_OK;
}

}  // namespace

extern "C" {

int lam_abi_version(void) { return LAM_ABI_VERSION; }

const char *lam_last_error(const lam_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lam_create(const lam_shape *sh, int mode, int num_buffers, lam_ctx **out) {
  if (!sh || !out) return LAM_ERR_ARG;
  *out = nullptr;
  if (mode != LAM_MODE_CORR_BEAM && mode != LAM_MODE_ADAPTIVE && mode != LAM_MODE_LAMBDA_GPU &&
      mode != LAM_MODE_PROJECTION)
    return LAM_ERR_ARG;
  if (sh->nr_polarizations != 2) return LAM_ERR_UNSUPPORTED;
  if (sh->nr_channels < 1 || sh->nr_fpga_sources < 1 || sh->nr_time_steps_per_packet < 1 ||
      sh->nr_receivers_per_packet < 1 || sh->nr_packets_for_correlation < 1 || sh->nr_beams < 1)
    return LAM_ERR_ARG;
  if (sh->nr_receivers != sh->nr_fpga_sources * sh->nr_receivers_per_packet) return LAM_ERR_ARG;
  if (sh->nr_receivers < 1 || sh->nr_receivers > kMaxReceivers) return LAM_ERR_UNSUPPORTED;
  if (num_buffers < 1 || num_buffers > 4) return LAM_ERR_ARG;
  const long T = (long)sh->nr_packets_for_correlation * sh->nr_time_steps_per_packet;
  // at most kMaxSplits splits of at most kMaxSamplesPerSplit samples each, so
  // every int32 partial sum stays exact
  if (T > 65536) return LAM_ERR_UNSUPPORTED;
  const int Bout = mode == LAM_MODE_ADAPTIVE ? 2 * sh->nr_beams : sh->nr_beams;
  if (Bout > 64 || (mode != LAM_MODE_ADAPTIVE && sh->nr_beams > 32)) return LAM_ERR_UNSUPPORTED;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return LAM_ERR_NO_DEVICE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return LAM_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return LAM_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return LAM_ERR_NO_DEVICE;

  lam_ctx *ctx = new lam_ctx();
  ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
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
  // SPATIAL_DISABLE_CUDA_GRAPH set (any value) forces eager launches, as in the
  // reference (lambda_corr_beam_only_gpu_pipeline.hpp:561); LAM_DISABLE_GRAPH=1 too
  if (std::getenv("SPATIAL_DISABLE_CUDA_GRAPH") != nullptr) ctx->use_graphs = false;
  if (const char *e = std::getenv("LAM_DISABLE_GRAPH")) ctx->use_graphs = ctx->use_graphs && e[0] != '1';

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
    if (hipEventCreateWithFlags(&b.eig_done, hipEventDisableTiming) != hipSuccess) rc = LAM_ERR_HIP;
    chk(dalloc(ctx, &b.d_samples, ctx->samples_bytes));
    chk(dalloc(ctx, &b.d_scales, ctx->scales_bytes));
    chk(dalloc(ctx, &b.d_beams, ctx->beams_bytes));
    if (mode == LAM_MODE_ADAPTIVE && ctx->NT > 10)
      chk(dalloc(ctx, &b.d_spectra, (size_t)ctx->C * 2 * ctx->Bout * (ctx->NT - 10) * sizeof(float)));
    chk(dalloc(ctx, &b.d_vis_part, (size_t)kMaxSplits * vis_elems * sizeof(int)));
    chk(dalloc(ctx, &b.d_err, sizeof(int)));
    if (hipHostMalloc(reinterpret_cast<void **>(&b.h_err), sizeof(int), hipHostMallocDefault) != hipSuccess)
      rc = LAM_ERR_HIP;
    else
      *b.h_err = 0;
    if (mode == LAM_MODE_LAMBDA_GPU) {  // eigenpairs of the four pol pairs
      chk(dalloc(ctx, &b.d_evals, (size_t)ctx->C * 4 * ctx->R * sizeof(float)));
      chk(dalloc(ctx, &b.d_evecs, (size_t)ctx->C * 4 * ctx->R * ctx->R * 2 * sizeof(float)));
    }
    if (mode == LAM_MODE_ADAPTIVE) {
      chk(dalloc(ctx, &b.d_wop2, (size_t)ctx->C * 2 * ctx->NMT2 * ctx->NK * 2 * 64 * sizeof(v8h)));
      chk(dalloc(ctx, &b.d_evals, (size_t)ctx->C * 2 * ctx->R * sizeof(float)));
      chk(dalloc(ctx, &b.d_evecs, (size_t)ctx->C * 2 * ctx->R * ctx->R * 2 * sizeof(float)));
      chk(dalloc(ctx, &b.d_counts, (size_t)ctx->C * 2 * sizeof(int)));
      chk(dalloc(ctx, &b.d_sweeps, (size_t)ctx->C * 2 * sizeof(int)));
      chk(dalloc(ctx, &b.d_fallback, (size_t)ctx->C * 2 * sizeof(int)));
      chk(dalloc(ctx, &b.d_nsig, (size_t)ctx->C * 2 * sizeof(int)));
      chk(dalloc(ctx, &b.d_orth, (size_t)ctx->C * 2 * sizeof(float)));
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
  if (mode == LAM_MODE_PROJECTION) {
    const size_t nprob = (size_t)ctx->C * 4;
    chk(dalloc(ctx, &ctx->d_proj_acc, nprob * ctx->R * ctx->R * 2 * sizeof(float)));
    chk(dalloc(ctx, &ctx->d_proj_evals, nprob * ctx->R * sizeof(float)));
    chk(dalloc(ctx, &ctx->d_proj_evecs, nprob * ctx->R * ctx->R * 2 * sizeof(float)));
  }
  if (mode == LAM_MODE_ADAPTIVE && ctx->NT == 64 && rc == LAM_OK) {
    std::vector<uint16_t> tab(spectra_table_bytes() / 2);
    spectra_table_host(tab.data());
    if (chk(dalloc(ctx, &ctx->d_spec_table, spectra_table_bytes())) &&
        hipMemcpy(ctx->d_spec_table, tab.data(), spectra_table_bytes(), hipMemcpyHostToDevice) != hipSuccess)
      rc = LAM_ERR_HIP;
  }
  chk(dalloc(ctx, &ctx->d_dump_f32, vis_elems * sizeof(float)));
  chk(dalloc(ctx, &ctx->d_dump_i64, vis_elems * sizeof(long long)));
  if (hipEventCreateWithFlags(&ctx->reset_done, hipEventDisableTiming) != hipSuccess) rc = LAM_ERR_HIP;
  if (rc == LAM_OK) {
    // unity gains and zero fine-delay phase (get_default_gains, pipeline/common.hpp:497-503)
    std::vector<float2> ones((size_t)ctx->C * 2 * ctx->R, make_float2(1.f, 0.f));
    if (hipMemcpy(ctx->d_gains, ones.data(), ones.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->d_phase, ones.data(), (size_t)ctx->C * ctx->R * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
      rc = LAM_ERR_HIP;
    if (rc == LAM_OK && hipEventRecord(ctx->reset_done, ctx->bufs[0].stream) != hipSuccess) rc = LAM_ERR_HIP;
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
    if (b.stream) (void)hipStreamSynchronize(b.stream);
  }
  for (auto &e : ctx->prof) {
    (void)hipEventDestroy(e.start);
    (void)hipEventDestroy(e.stop);
  }
  for (auto &b : ctx->bufs) {
    destroy_graphs(b);
    if (b.done) (void)hipEventDestroy(b.done);
    if (b.eig_done) (void)hipEventDestroy(b.eig_done);
    (void)hipFree(b.d_samples);
    (void)hipFree(b.d_scales);
    (void)hipFree(b.d_beams);
    (void)hipFree(b.d_spectra);
    (void)hipFree(b.d_vis_part);
    (void)hipFree(b.d_err);
    if (b.h_err) (void)hipHostFree(b.h_err);
    (void)hipFree(b.d_wop2);
    (void)hipFree(b.d_evals);
    (void)hipFree(b.d_evecs);
    (void)hipFree(b.d_counts);
    (void)hipFree(b.d_sweeps);
    (void)hipFree(b.d_fallback);
    (void)hipFree(b.d_nsig);
    (void)hipFree(b.d_orth);
    if (b.stream) (void)hipStreamDestroy(b.stream);
  }
  (void)hipFree(ctx->d_vis_acc);
  (void)hipFree(ctx->d_weights);
  (void)hipFree(ctx->d_gains);
  (void)hipFree(ctx->d_phase);
  (void)hipFree(ctx->d_delays);
  (void)hipFree(ctx->d_recv_perm);
  (void)hipFree(ctx->d_pol_perm);
  (void)hipFree(ctx->d_wop);
  (void)hipFree(ctx->d_fixed_k);
  (void)hipFree(ctx->d_spec_table);
  (void)hipFree(ctx->d_proj_acc);
  (void)hipFree(ctx->d_proj_evals);
  (void)hipFree(ctx->d_proj_evecs);
  (void)hipFree(ctx->d_dump_f32);
  (void)hipFree(ctx->d_dump_i64);
  free_learn(ctx->lr);
  free_subarrays(ctx);
  if (ctx->reset_done) (void)hipEventDestroy(ctx->reset_done);
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

int lam_set_spectra(lam_ctx *ctx, int enable) {
  if (!ctx) return LAM_ERR_ARG;
  if (enable && (ctx->mode != LAM_MODE_ADAPTIVE || ctx->NT != 64 || ctx->T % 64 != 0 || !ctx->bufs[0].d_spectra ||
                 !ctx->d_spec_table))
    return fail(ctx, LAM_ERR_ARG, "spectra need the adaptive mode and 64 time steps per packet");
  if (int r = sync_all(ctx)) return r;
  ctx->spectra = enable != 0;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_set_whitening(lam_ctx *ctx, int enable) {
  if (!ctx) return LAM_ERR_ARG;
  if (ctx->mode != LAM_MODE_ADAPTIVE) return fail(ctx, LAM_ERR_STATE, "whitening applies to the adaptive mode");
  if (!enable && ctx->eig_method != LAM_EIG_PLAIN)
    return fail(ctx, LAM_ERR_STATE, "the warm starts work on the whitened covariance: set the plain solver first");
  if (!enable && ctx->sub_G > 0)
    return fail(ctx, LAM_ERR_STATE, "the subarray eigen stage always whitens: clear the subarrays first");
  if (int r = sync_all(ctx)) return r;
  ctx->whiten = enable ? 1 : 0;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_set_bandpass(lam_ctx *ctx, int downsample_factor) {
  if (!ctx || downsample_factor < 0) return LAM_ERR_ARG;
  if (ctx->mode != LAM_MODE_LAMBDA_GPU)  // (the adaptive spectra share the landing buffers)
    return downsample_factor == 0 ? LAM_OK : fail(ctx, LAM_ERR_ARG, "bandpass needs the LambdaGPU mode");
  if (downsample_factor > 0 &&
      (ctx->mode != LAM_MODE_LAMBDA_GPU || ctx->T > 8192 || (ctx->T & (ctx->T - 1)) || ctx->T % downsample_factor))
    return fail(ctx, LAM_ERR_ARG,
                "bandpass needs the LambdaGPU mode, a power-of-two block length <= 8192 and a factor dividing it");
  if (int r = sync_all(ctx)) return r;
  for (auto &b : ctx->bufs) {
    (void)hipFree(b.d_spectra);
    b.d_spectra = nullptr;
  }
  ctx->bandpass = 0;
  ctx->graphs_dirty = true;
  if (downsample_factor == 0) return LAM_OK;
  for (auto &b : ctx->bufs)
    if (int r = dalloc(ctx, &b.d_spectra, (size_t)ctx->C * 2 * ctx->B * (ctx->T / downsample_factor) * sizeof(float)))
      return r;
  ctx->bandpass = downsample_factor;
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

int lam_set_weight_precision(lam_ctx *ctx, int precision) {
  if (!ctx || precision < LAM_WPREC_F16_HILO || precision > LAM_WPREC_BF16) return LAM_ERR_ARG;
  if (int r = sync_all(ctx)) return r;
  ctx->wprec = precision;
  ctx->graphs_dirty = true;
  return update_wop(ctx);
}

int lam_set_eig_tuning(lam_ctx *ctx, int plain_threads) {
  auto ok = [](int t) { return t == 64 || t == 128 || t == 256; };
  if (!ctx || !ok(plain_threads)) return LAM_ERR_ARG;
  ctx->eig_plain_threads = plain_threads;
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
    (void)hipEventDestroy(e.start);
    (void)hipEventDestroy(e.stop);
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

static int execute_block(lam_ctx *ctx, const lam_block_io *io, bool &release_queued) {
  if (io->spectra_dst && !ctx->spectra && !ctx->bandpass)
    return fail(ctx, LAM_ERR_ARG, "spectra_dst given but spectra are off");
  Buf &b = ctx->bufs[ctx->cur];
  hipStream_t s = b.stream;
  const void *smp = io->samples, *scl = io->scales;
  if (!io->inputs_on_device) {
    // H2D of the block, then release the host buffer (pipeline/common.hpp:517-528)
    HIPC(ctx, hipMemcpyAsync(b.d_samples, io->samples, ctx->samples_bytes, hipMemcpyHostToDevice, s));
    HIPC(ctx, hipMemcpyAsync(b.d_scales, io->scales, ctx->scales_bytes, hipMemcpyHostToDevice, s));
    if (io->release_fn) {
      HIPC(ctx, hipLaunchHostFunc(s, io->release_fn, io->release_user));
      release_queued = true;
    }
    smp = b.d_samples;
    scl = b.d_scales;
  }
  // every eigenpair is formed only for a block whose eigen outputs are requested
  ctx->eig_full = (io->eigenvalues_dst || io->eigenvectors_dst) ? 1 : 0;
  // never accumulate into a visibility buffer that a dump has not reset yet
  HIPC(ctx, hipStreamWaitEvent(s, ctx->reset_done, 0));
  HIPC(ctx, hipMemsetAsync(b.d_err, 0, sizeof(int), s));
  if (int r = run_block(ctx, b, smp, scl)) return r;
  if (io->inputs_on_device && io->release_fn) {
    HIPC(ctx, hipLaunchHostFunc(s, io->release_fn, io->release_user));
    release_queued = true;
  }
  if (io->beams_dst) {
    HIPC(ctx, hipMemcpyAsync(io->beams_dst, b.d_beams, ctx->beams_bytes, hipMemcpyDefault, s));
    if (io->beams_done_fn) HIPC(ctx, hipLaunchHostFunc(s, io->beams_done_fn, io->beams_done_user));
  }
  if (ctx->mode == LAM_MODE_LAMBDA_GPU && io->spectra_dst) {
    HIPC(ctx, hipMemcpyAsync(io->spectra_dst, b.d_spectra,
                             (size_t)ctx->C * 2 * ctx->B * (ctx->T / ctx->bandpass) * sizeof(float),
                             hipMemcpyDefault, s));
    if (io->spectra_done_fn) HIPC(ctx, hipLaunchHostFunc(s, io->spectra_done_fn, io->spectra_done_user));
  }
  if (ctx->mode == LAM_MODE_LAMBDA_GPU && (io->eigenvalues_dst || io->eigenvectors_dst)) {
    if (io->eigenvalues_dst)
      HIPC(ctx, hipMemcpyAsync(io->eigenvalues_dst, b.d_evals, (size_t)ctx->C * 4 * ctx->R * sizeof(float),
                               hipMemcpyDefault, s));
    if (io->eigenvectors_dst)
      HIPC(ctx, hipMemcpyAsync(io->eigenvectors_dst, b.d_evecs,
                               (size_t)ctx->C * 4 * ctx->R * ctx->R * 2 * sizeof(float), hipMemcpyDefault, s));
    if (io->eigen_done_fn) HIPC(ctx, hipLaunchHostFunc(s, io->eigen_done_fn, io->eigen_done_user));
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
    if (io->spectra_dst) {
      HIPC(ctx, hipMemcpyAsync(io->spectra_dst, b.d_spectra,
                               (size_t)ctx->C * 2 * ctx->Bout * (ctx->NT - 10) * sizeof(float),
                               hipMemcpyDefault, s));
      if (io->spectra_done_fn) HIPC(ctx, hipLaunchHostFunc(s, io->spectra_done_fn, io->spectra_done_user));
    }
  }
  HIPC(ctx, hipMemcpyAsync(b.h_err, b.d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPC(ctx, hipEventRecord(b.done, s));
  b.err_pending = true;
  b.err_block = ctx->blocks_run++;
  if (ctx->mode == LAM_MODE_PROJECTION) ++ctx->proj_runs;
  ctx->last_buf = ctx->cur;
  ctx->blocks_acc += 1;
  ctx->cur = (ctx->cur + 1) % ctx->nbuf;
  return LAM_OK;
}

int lam_execute(lam_ctx *ctx, const lam_block_io *io) {
  if (!ctx || !io) return LAM_ERR_ARG;
  bool release_queued = false;
  int rc = LAM_OK;
  if (!io->samples || !io->scales) rc = fail(ctx, LAM_ERR_ARG, "execute: samples and scales are required");
  // a range error of an earlier, completed block fails this call
  if (rc == LAM_OK) rc = check_range(ctx, false);
  if (rc == LAM_OK) rc = execute_block(ctx, io, release_queued);
  if (rc != LAM_OK && io->release_fn && !release_queued) {
    // the host input is released exactly once whether or not the block ran:
    // once nothing on the stream reads it any more
    Buf &b = ctx->bufs[ctx->cur];
    (void)hipStreamSynchronize(b.stream);
    (void)hipGetLastError();
    io->release_fn(io->release_user);
  }
  return rc;
}

static int dump_common(lam_ctx *ctx, float *f32, int64_t *i64, lam_host_fn fn, void *user) {
  // a block with a range error reached this accumulator: fail the dump
  if (int r = check_range(ctx, false)) return r;
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

int lam_set_projection(lam_ctx *ctx, int nr_signal_eigenvectors, int nr_runs_to_average) {
  if (!ctx) return LAM_ERR_ARG;
  if (ctx->mode != LAM_MODE_PROJECTION) return fail(ctx, LAM_ERR_STATE, "projection settings need the projection mode");
  if (nr_signal_eigenvectors < 1 || nr_signal_eigenvectors > ctx->R || nr_runs_to_average < 1)
    return fail(ctx, LAM_ERR_ARG, "need 1 <= nr_signal_eigenvectors <= %d and nr_runs_to_average >= 1", ctx->R);
  if (int r = sync_all(ctx)) return r;
  ctx->proj_k = nr_signal_eigenvectors;
  ctx->proj_avg = nr_runs_to_average;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_projection_runs(const lam_ctx *ctx) { return ctx ? ctx->proj_runs : 0; }

int lam_dump_projection(lam_ctx *ctx, float *evals_dst, float *evecs_dst, lam_host_fn fn, void *user) {
  if (!ctx) return LAM_ERR_ARG;
  if (ctx->mode != LAM_MODE_PROJECTION) return fail(ctx, LAM_ERR_STATE, "no projection accumulator in this mode");
  if (ctx->proj_runs == 0) return fail(ctx, LAM_ERR_STATE, "no block integrated since the last projection dump");
  if (int r = check_range(ctx, false)) return r;
  hipStream_t s = ctx->bufs[0].stream;
  for (auto &b : ctx->bufs) HIPC(ctx, hipStreamWaitEvent(s, b.done, 0));
  const size_t nprob = (size_t)ctx->C * 4;
  HIPC(ctx, launch_proj_dump(ctx->d_proj_acc, ctx->C, ctx->R, ctx->proj_runs, ctx->max_sweeps, ctx->tol,
                             ctx->d_proj_evals, ctx->d_proj_evecs, s));
  HIPC(ctx, hipEventRecord(ctx->reset_done, s));  // later blocks add into the cleared accumulator
  if (evals_dst)
    HIPC(ctx, hipMemcpyAsync(evals_dst, ctx->d_proj_evals, nprob * ctx->R * sizeof(float), hipMemcpyDefault, s));
  if (evecs_dst)
    HIPC(ctx, hipMemcpyAsync(evecs_dst, ctx->d_proj_evecs, nprob * ctx->R * ctx->R * 2 * sizeof(float),
                             hipMemcpyDefault, s));
  if (fn) HIPC(ctx, hipLaunchHostFunc(s, fn, user));
  ctx->proj_runs = 0;
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
  return check_range(ctx, true);
}

long long lam_graph_replays(const lam_ctx *ctx) { return ctx ? ctx->graph_replays : -1; }

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

int lam_autotune(lam_ctx *ctx, const void *const *samples, const void *const *scales, int nblocks, int iters,
                 int32_t *result) {
  if (!ctx || !samples || !scales || nblocks < 1 || iters < 1) return LAM_ERR_ARG;
  for (int i = 0; i < nblocks; ++i)
    if (!samples[i] || !scales[i]) return LAM_ERR_ARG;
  if (int r = sync_all(ctx)) return r;
  const bool prof_was = ctx->profiling;
  double d0[LAM_K_COUNT] = {0};
  int d1[LAM_K_COUNT] = {0};
  if (int r = drain_prof(ctx, d0, d1)) return r;  // keep the caller's pending timings out
  for (int k = 0; k < LAM_K_COUNT; ++k) {
    ctx->prof_ms[k] += d0[k];
    ctx->prof_n[k] += d1[k];
  }
  ctx->profiling = true;
  // tuning blocks are no training examples: capture is suspended meanwhile
  Learn &L = ctx->lr;
  const int cap_max_was = L.cap_max;
  L.cap_max = L.cap_count;
  int rc = LAM_OK;
  // With several buffers consecutive blocks overlap, and the eigen stage
  // shares the GPU with the voltage kernels of neighbouring blocks: the
  // choices are then judged by streaming time per block, over at least 16
  // blocks so that the overlap reaches its steady state (measured on 384
  // channels: 1024 samples per workgroup gives the shortest serial kernels
  // but the longest streaming block, 0.71 against 0.61 ms at 2048-3072).
  const bool stream = ctx->nbuf > 1;
  const int stream_iters = std::max(iters, 16);
  constexpr int kRounds = 3;
  // warm-up: after idle the GPU runs the first ~50 blocks slower
  {
    double t = 0.0;
    rc = time_stream(ctx, samples, scales, nblocks, std::max(1, (48 + nblocks - 1) / nblocks), t);
  }
  // 1. time samples per voltage workgroup: more splits help the voltage
  //    kernels but multiply the partial visibilities that the eigen stage,
  //    the features and the accumulator read, so the whole sequence is timed
  if (rc == LAM_OK) {
    const int Tr = (ctx->T + kChunk - 1) / kChunk * kChunk;
    std::vector<int> cands;
    for (int cand : {0, 1024, 2048, 3072, 4096, 8192})
      if (cand <= Tr) cands.push_back(cand);
    int chosen = ctx->samples_per_wg;
    rc = pick_by_time(
        cands, ctx->samples_per_wg, kRounds, [&](int v) { ctx->samples_per_wg = v; },
        [&](double &t) {
          return stream ? time_stream(ctx, samples, scales, nblocks, stream_iters, t)
                        : time_sequence(ctx, samples, scales, nblocks, iters,
                                        {LAM_K_CORRBEAM, LAM_K_CORR, LAM_K_BEAM2B, LAM_K_EIG, LAM_K_FEAT, LAM_K_MLP,
                                         LAM_K_SUBARRAY},
                                        t);
        },
        chosen);
  }
  // 2. plain eigen solve: workgroup width
  if (rc == LAM_OK && ctx->mode == LAM_MODE_ADAPTIVE) {
    const int method = ctx->eig_method;
    ctx->eig_method = LAM_EIG_PLAIN;
    int chosen = ctx->eig_plain_threads;
    rc = pick_by_time(
        {64, 128, 256}, ctx->eig_plain_threads, kRounds,
        [&](int v) { ctx->eig_plain_threads = v; },
        [&](double &t) {
          return stream ? time_stream(ctx, samples, scales, nblocks, stream_iters, t)
                        : time_sequence(ctx, samples, scales, nblocks, iters, {LAM_K_EIG}, t);
        },
        chosen);
    ctx->eig_method = method;
    // 3. warm start: refinement steps (rejections fall back to the plain solver
    //    inside the warm launch and show up in its time).  Judged by the
    //    eigen stage's own kernel time even with several buffers: the step
    //    count moves it by ~10 us, which the streaming time per block (bound
    //    by the voltage passes of neighbouring blocks) does not resolve.
    if (rc == LAM_OK && method != LAM_EIG_PLAIN && ctx->lr.ready) {
      rc = pick_by_time(
          {1, 2, 3}, ctx->refine_sweeps, kRounds, [&](int v) { ctx->refine_sweeps = v; },
          [&](double &t) {
            return time_sequence(ctx, samples, scales, nblocks, iters, {LAM_K_EIG, LAM_K_FEAT, LAM_K_MLP}, t);
          },
          chosen);
    }
  }
  L.cap_max = cap_max_was;
  ctx->profiling = prof_was;
  ctx->graphs_dirty = true;
  // the tuning blocks must not reach the visibility accumulator
  if (rc == LAM_OK) {
    hipStream_t st = ctx->bufs[0].stream;
    if (launch_zero(ctx->d_vis_acc, (size_t)ctx->C * ctx->NBL * 8 * sizeof(long long), st) != hipSuccess ||
        (ctx->d_proj_acc &&
         launch_zero(ctx->d_proj_acc, (size_t)ctx->C * 4 * ctx->R * ctx->R * 2 * sizeof(float), st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = fail(ctx, LAM_ERR_HIP, "autotune: clearing the accumulator failed");
    ctx->proj_runs = 0;
  }
  // range errors of the tuning blocks (run outside lam_execute, whose flag
  // copies they do not have)
  for (auto &b : ctx->bufs) {
    int flag = 0;
    if (hipMemcpy(&flag, b.d_err, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
      if (rc == LAM_OK) rc = fail(ctx, LAM_ERR_HIP, "autotune: reading the range flag failed");
      continue;
    }
    if (flag) {
      (void)hipMemset(b.d_err, 0, sizeof(int));
      if (rc == LAM_OK)
        rc = fail(ctx, LAM_ERR_RANGE, "autotune: a tuning block has samples outside the int8 range after scaling");
    }
  }
  if (result) {
    result[0] = ctx->samples_per_wg;
    result[1] = ctx->eig_plain_threads;
    result[2] = ctx->refine_sweeps;
  }
  return rc;
}

int lam_set_subarrays(lam_ctx *ctx, int ngroups, const uint8_t *mask, const int32_t *k) {
  if (!ctx || ngroups < 0 || (ngroups > 0 && (!mask || !k))) return LAM_ERR_ARG;
  if (ctx->mode != LAM_MODE_ADAPTIVE) return fail(ctx, LAM_ERR_STATE, "subarrays need the adaptive mode");
  if (ngroups > 0 && !ctx->whiten)
    return fail(ctx, LAM_ERR_STATE, "the subarray eigen stage whitens: turn whitening back on first");
  if (ngroups > 0 && (2 + ngroups) * ctx->B > 64)
    return fail(ctx, LAM_ERR_UNSUPPORTED, "(2 + groups) x beams = %d output beams per (channel, pol); at most 64",
                (2 + ngroups) * ctx->B);
  if (int r = sync_all(ctx)) return r;
  free_subarrays(ctx);
  ctx->graphs_dirty = true;
  for (auto &b : ctx->bufs) destroy_graphs(b);
  if (ngroups == 0) return size_beam_outputs(ctx, 2 * ctx->B);
  const int C = ctx->C, R = ctx->R, G = ngroups, np = C * 2 * G;
  // problem order: largest subarray first (the persistent workgroups take them in turn)
  std::vector<int> size(np), order(np);
  for (int c = 0; c < C; ++c)
    for (int p = 0; p < 2; ++p)
      for (int g = 0; g < G; ++g) {
        int n = 0;
        for (int r = 0; r < R; ++r) n += mask[((size_t)c * G + g) * R + r] != 0;
        size[(c * 2 + p) * G + g] = n;
      }
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return size[a] > size[b]; });
  // one queue with the 2C full-array problems (items 0..2C-1, size R) and the
  // subarray problems (items 2C + q), largest first
  std::vector<int> qsize(2 * C + np), qorder(2 * C + np);
  for (int i = 0; i < 2 * C; ++i) qsize[i] = R;
  for (int q = 0; q < np; ++q) qsize[2 * C + q] = size[q];
  std::iota(qorder.begin(), qorder.end(), 0);
  std::stable_sort(qorder.begin(), qorder.end(), [&](int a, int b) { return qsize[a] > qsize[b]; });
  int rc = LAM_OK;
  auto chk = [&](int r) {
    if (r != LAM_OK && rc == LAM_OK) rc = r;
  };
  chk(dalloc(ctx, &ctx->d_sub_mask, (size_t)C * G * R));
  chk(dalloc(ctx, &ctx->d_sub_k, (size_t)np * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_sub_order, (size_t)np * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_sub_counter, sizeof(int)));
  chk(dalloc(ctx, &ctx->d_sub_evals, (size_t)np * R * sizeof(float)));
  chk(dalloc(ctx, &ctx->d_sub_evecs, (size_t)np * R * R * 2 * sizeof(float)));
  chk(dalloc(ctx, &ctx->d_sub_proj, (size_t)np * R * R * 2 * sizeof(float)));
  chk(dalloc(ctx, &ctx->d_sub_counts, (size_t)np * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_sub_sweeps, (size_t)np * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_sub_sizes, (size_t)np * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_queue_order, qorder.size() * sizeof(int)));
  chk(dalloc(ctx, &ctx->d_queue_counter, sizeof(int)));
  if (rc == LAM_OK) rc = size_beam_outputs(ctx, (2 + G) * ctx->B);
  if (rc != LAM_OK) {
    free_subarrays(ctx);
    return rc;
  }
  HIPC(ctx, hipMemcpy(ctx->d_queue_order, qorder.data(), qorder.size() * sizeof(int), hipMemcpyHostToDevice));
  ctx->queue_n = (int)qorder.size();
  HIPC(ctx, hipMemcpy(ctx->d_sub_mask, mask, (size_t)C * G * R, hipMemcpyHostToDevice));
  HIPC(ctx, hipMemcpy(ctx->d_sub_k, k, (size_t)np * sizeof(int), hipMemcpyHostToDevice));
  HIPC(ctx, hipMemcpy(ctx->d_sub_order, order.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice));
  ctx->sub_G = G;
  ctx->sub_max_n = std::max(1, *std::max_element(size.begin(), size.end()));
  return LAM_OK;
}

int lam_set_persistent_workgroups(lam_ctx *ctx, int max_workgroups) {
  if (!ctx || max_workgroups < 0) return LAM_ERR_ARG;
  ctx->persistent_wgs = max_workgroups;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_subarray_results(lam_ctx *ctx, float *evals, float *evecs, float *proj, int32_t *counts,
                         int32_t *sizes, int32_t *sweeps) {
  if (!ctx) return LAM_ERR_ARG;
  if (ctx->sub_G == 0) return fail(ctx, LAM_ERR_STATE, "no subarrays set");
  if (int r = sync_all(ctx)) return r;
  const size_t np = (size_t)ctx->C * 2 * ctx->sub_G, R = ctx->R;
  if (evals) HIPC(ctx, hipMemcpy(evals, ctx->d_sub_evals, np * R * sizeof(float), hipMemcpyDeviceToHost));
  if (evecs) HIPC(ctx, hipMemcpy(evecs, ctx->d_sub_evecs, np * R * R * 2 * sizeof(float), hipMemcpyDeviceToHost));
  if (proj) HIPC(ctx, hipMemcpy(proj, ctx->d_sub_proj, np * R * R * 2 * sizeof(float), hipMemcpyDeviceToHost));
  if (counts) HIPC(ctx, hipMemcpy(counts, ctx->d_sub_counts, np * sizeof(int), hipMemcpyDeviceToHost));
  if (sizes) HIPC(ctx, hipMemcpy(sizes, ctx->d_sub_sizes, np * sizeof(int), hipMemcpyDeviceToHost));
  if (sweeps) HIPC(ctx, hipMemcpy(sweeps, ctx->d_sub_sweeps, np * sizeof(int), hipMemcpyDeviceToHost));
  return LAM_OK;
}

int lam_set_eig_method(lam_ctx *ctx, int method, int refine_sweeps, int power, float residual_tol) {
  if (!ctx) return LAM_ERR_ARG;
  if (method < LAM_EIG_PLAIN || method > LAM_EIG_PREVIOUS || refine_sweeps < 0 || refine_sweeps > 16 ||
      !(residual_tol >= 0.0f))
    return fail(ctx, LAM_ERR_ARG, "lam_set_eig_method: bad argument");
  if (ctx->mode != LAM_MODE_ADAPTIVE) return fail(ctx, LAM_ERR_STATE, "eigen method needs the adaptive mode");
  if (method != LAM_EIG_PLAIN && !ctx->whiten)
    return fail(ctx, LAM_ERR_STATE, "the warm starts work on the whitened covariance: whitening is off");
  if (method != LAM_EIG_PLAIN)
    if (int r = ensure_learn(ctx, 4, 256, 1, false)) return r;
  if (int r = sync_all(ctx)) return r;
  ctx->eig_method = method;
  ctx->refine_sweeps = refine_sweeps;
  ctx->power = power ? 1 : 0;
  ctx->res_tol = residual_tol;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_eig_stats(lam_ctx *ctx, int32_t *sweeps, int32_t *fallback, float *orthogonality) {
  if (!ctx) return LAM_ERR_ARG;
  if (ctx->mode != LAM_MODE_ADAPTIVE || ctx->last_buf < 0) return fail(ctx, LAM_ERR_STATE, "no adaptive block run yet");
  if (int r = sync_all(ctx)) return r;
  Buf &b = ctx->bufs[ctx->last_buf];
  const size_t nb = (size_t)ctx->C * 2 * sizeof(int);
  if (sweeps) HIPC(ctx, hipMemcpy(sweeps, b.d_sweeps, nb, hipMemcpyDeviceToHost));
  if (fallback) HIPC(ctx, hipMemcpy(fallback, b.d_fallback, nb, hipMemcpyDeviceToHost));
  if (orthogonality) HIPC(ctx, hipMemcpy(orthogonality, b.d_orth, nb, hipMemcpyDeviceToHost));
  return LAM_OK;
}

int lam_learn_init(lam_ctx *ctx, int warm_vectors, int hidden, uint64_t seed) {
  if (!ctx) return LAM_ERR_ARG;
  if (ctx->mode != LAM_MODE_ADAPTIVE) return fail(ctx, LAM_ERR_STATE, "learning needs the adaptive mode");
  if (ctx->R < 2 * warm_vectors) return fail(ctx, LAM_ERR_ARG, "need receivers >= 2 * warm_vectors");
  return ensure_learn(ctx, warm_vectors, hidden, seed, true);
}

int lam_learn_dims(lam_ctx *ctx, int32_t *dims, int64_t *params) {
  if (!ctx) return LAM_ERR_ARG;
  if (int r = ensure_learn(ctx, 4, 256, 1, false)) return r;
  const Learn &L = ctx->lr;
  if (dims) {
    const int32_t d[7] = {L.KW, L.n, L.Dx, L.Dp, L.H, L.Dy, L.Dyp};
    std::memcpy(dims, d, sizeof d);
  }
  if (params) *params = L.P;
  return LAM_OK;
}

int lam_learn_capture(lam_ctx *ctx, int max_blocks) {
  if (!ctx || max_blocks < 0) return LAM_ERR_ARG;
  if (int r = ensure_learn(ctx, 4, 256, 1, false)) return r;
  if (int r = sync_all(ctx)) return r;
  Learn &L = ctx->lr;
  if (L.d_dsX) (void)hipFree(L.d_dsX);
  if (L.d_dsU) (void)hipFree(L.d_dsU);
  if (L.d_dsK) (void)hipFree(L.d_dsK);
  L.d_dsX = nullptr;
  L.d_dsU = nullptr;
  L.d_dsK = nullptr;
  L.cap_max = L.cap_count = 0;
  if (max_blocks == 0) return LAM_OK;
  const size_t M = (size_t)ctx->C * 2;
  if (int r = dalloc(ctx, &L.d_dsX, max_blocks * M * L.Dp * sizeof(uint16_t))) return r;
  if (int r = dalloc(ctx, &L.d_dsU, max_blocks * M * L.KW * L.n * sizeof(float2))) return r;
  if (int r = dalloc(ctx, &L.d_dsK, max_blocks * M * sizeof(int))) return r;
  L.cap_max = max_blocks;
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_learn_captured(lam_ctx *ctx) { return ctx ? ctx->lr.cap_count : 0; }

int lam_learn_last(lam_ctx *ctx, uint16_t *x, float *y) {
  if (!ctx) return LAM_ERR_ARG;
  Learn &L = ctx->lr;
  if (!L.ready) return fail(ctx, LAM_ERR_STATE, "learn_last: no network");
  if (int r = sync_all(ctx)) return r;
  const size_t M = (size_t)ctx->C * 2;
  if (x) HIPC(ctx, hipMemcpy(x, L.d_X, M * L.Dp * sizeof(uint16_t), hipMemcpyDeviceToHost));
  if (y) HIPC(ctx, hipMemcpy(y, L.d_Y, M * L.Dyp * sizeof(float), hipMemcpyDeviceToHost));
  return LAM_OK;
}

int lam_learn_sample(lam_ctx *ctx, int block, uint16_t *x, float *labels, int32_t *k) {
  if (!ctx) return LAM_ERR_ARG;
  Learn &L = ctx->lr;
  if (!L.ready || block < 0 || block >= L.cap_count)
    return fail(ctx, LAM_ERR_ARG, "learn_sample: block %d not captured", block);
  if (int r = sync_all(ctx)) return r;
  const size_t M = (size_t)ctx->C * 2, nu = (size_t)L.KW * ctx->R;
  if (x)
    HIPC(ctx, hipMemcpy(x, L.d_dsX + (size_t)block * M * L.Dp, M * L.Dp * sizeof(uint16_t), hipMemcpyDeviceToHost));
  if (labels)
    HIPC(ctx, hipMemcpy(labels, L.d_dsU + (size_t)block * M * nu, M * nu * sizeof(float2), hipMemcpyDeviceToHost));
  if (k) HIPC(ctx, hipMemcpy(k, L.d_dsK + (size_t)block * M, M * sizeof(int), hipMemcpyDeviceToHost));
  return LAM_OK;
}

int lam_learn_train(lam_ctx *ctx, int epochs, int batch, float lr, uint64_t seed, float *loss_per_epoch) {
  if (!ctx || epochs < 0 || batch < 1 || !(lr > 0.0f)) return LAM_ERR_ARG;
  Learn &L = ctx->lr;
  if (!L.ready || L.cap_count == 0) return fail(ctx, LAM_ERR_STATE, "no captured training blocks");
  if (int r = sync_all(ctx)) return r;
  if (int r = ensure_batch(ctx, batch)) return r;
  const long N = (long)L.cap_count * ctx->C * 2;
  if (L.idx_cap < N) {
    if (L.d_idx) (void)hipFree(L.d_idx);
    L.d_idx = nullptr;
    L.idx_cap = 0;
    if (int r = dalloc(ctx, &L.d_idx, N * sizeof(int))) return r;
    L.idx_cap = N;
  }
  hipStream_t s = ctx->bufs[0].stream;
  std::vector<int> perm(N);
  std::mt19937_64 gen(seed);
  const int nu = L.KW * L.n;
  for (int ep = 0; ep < epochs; ++ep) {
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), gen);
    HIPC(ctx, hipMemcpyAsync(L.d_idx, perm.data(), N * sizeof(int), hipMemcpyHostToDevice, s));
    HIPC(ctx, launch_zero(L.d_loss, sizeof(float), s));
    int steps = 0;
    for (long b0 = 0; b0 < N; b0 += batch) {
      const int nb = (int)std::min<long>(batch, N - b0);
      HIPC(ctx, launch_gather(L.d_idx + b0, nb, L.d_dsX, L.Dp, L.d_dsU, nu, L.d_dsK, L.d_bX, L.d_bU, L.d_bK, s));
      if (int r = forward_backward(ctx, nb, s)) return r;
      L.adam_t += 1;
      HIPC(ctx, launch_adam(L.d_prm, L.d_grad, L.d_m1, L.d_m2, L.P, lr, 0.9f, 0.999f, 1e-8f, L.adam_t, 0,
                            (long)L.H * L.Dp, L.d_W1bf, L.off_W2, (long)L.Dyp * L.H, L.d_W2bf, s));
      ++steps;
    }
    float tot = 0.0f;
    HIPC(ctx, hipMemcpyAsync(&tot, L.d_loss, sizeof(float), hipMemcpyDeviceToHost, s));
    HIPC(ctx, hipStreamSynchronize(s));
    if (loss_per_epoch) loss_per_epoch[ep] = tot / (float)std::max(1, steps);
  }
  ctx->graphs_dirty = true;
  return LAM_OK;
}

int lam_learn_params(lam_ctx *ctx, float *params, int64_t count, int set) {
  if (!ctx || !params) return LAM_ERR_ARG;
  if (int r = ensure_learn(ctx, 4, 256, 1, false)) return r;
  Learn &L = ctx->lr;
  if (count != L.P) return fail(ctx, LAM_ERR_ARG, "parameter count %lld != %ld", (long long)count, L.P);
  if (int r = sync_all(ctx)) return r;
  if (!set) {
    HIPC(ctx, hipMemcpy(params, L.d_prm, L.P * sizeof(float), hipMemcpyDeviceToHost));
    return LAM_OK;
  }
  HIPC(ctx, hipMemcpy(L.d_prm, params, L.P * sizeof(float), hipMemcpyHostToDevice));
  hipStream_t st = ctx->bufs[0].stream;
  HIPC(ctx, launch_to_bf16(L.d_prm, (long)L.H * L.Dp, L.d_W1bf, st));
  HIPC(ctx, launch_to_bf16(L.d_prm + L.off_W2, (long)L.Dyp * L.H, L.d_W2bf, st));
  HIPC(ctx, hipStreamSynchronize(st));
  L.adam_t = 0;
  HIPC(ctx, hipMemset(L.d_m1, 0, L.P * sizeof(float)));
  HIPC(ctx, hipMemset(L.d_m2, 0, L.P * sizeof(float)));
  return LAM_OK;
}

int lam_learn_grad(lam_ctx *ctx, const uint16_t *x, const float *labels, const int32_t *k, int m,
                   float *loss, float *grad) {
  if (!ctx || !x || !labels || !k || m < 1) return LAM_ERR_ARG;
  if (int r = ensure_learn(ctx, 4, 256, 1, false)) return r;
  if (int r = sync_all(ctx)) return r;
  if (int r = ensure_batch(ctx, m)) return r;
  Learn &L = ctx->lr;
  hipStream_t s = ctx->bufs[0].stream;
  HIPC(ctx, hipMemcpy(L.d_bX, x, (size_t)m * L.Dp * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIPC(ctx, hipMemcpy(L.d_bU, labels, (size_t)m * L.KW * L.n * sizeof(float2), hipMemcpyHostToDevice));
  HIPC(ctx, hipMemcpy(L.d_bK, k, (size_t)m * sizeof(int), hipMemcpyHostToDevice));
  HIPC(ctx, launch_zero(L.d_loss, sizeof(float), s));
  if (int r = forward_backward(ctx, m, s)) return r;
  HIPC(ctx, hipStreamSynchronize(s));
  if (loss) HIPC(ctx, hipMemcpy(loss, L.d_loss, sizeof(float), hipMemcpyDeviceToHost));
  if (grad) HIPC(ctx, hipMemcpy(grad, L.d_grad, L.P * sizeof(float), hipMemcpyDeviceToHost));
  return LAM_OK;
}

int lam_synth_block(const lam_shape *sh, uint64_t seed, int32_t block_index, float sigma, int32_t n_sources,
                    const float *d_tones, const float *d_steer, const float *d_gain, int8_t *d_samples,
                    int16_t *d_scales, void *stream) {
  if (!sh || !d_gain || !d_samples || !d_scales || n_sources < 0 || n_sources > kGenMaxSources ||
      (n_sources > 0 && (!d_tones || !d_steer)))
    return LAM_ERR_ARG;
  if (sh->nr_receivers != sh->nr_fpga_sources * sh->nr_receivers_per_packet) return LAM_ERR_ARG;
  GenParams g{};
  g.C = sh->nr_channels;
  g.F = sh->nr_fpga_sources;
  g.NT = sh->nr_time_steps_per_packet;
  g.R = sh->nr_receivers;
  g.RPP = sh->nr_receivers_per_packet;
  g.NP = sh->nr_packets_for_correlation;
  g.S = n_sources;
  g.seed = seed;
  g.block_index = block_index;
  g.sigma = sigma;
  g.tones = d_tones;
  g.steer = reinterpret_cast<const float2 *>(d_steer);
  g.gain = d_gain;
  g.samples = d_samples;
  g.scales = d_scales;
  if (launch_synth(g, static_cast<hipStream_t>(stream)) != hipSuccess) return LAM_ERR_HIP;
  // NULL stream: return only when the block is written (the pipeline's own
  // streams are non-blocking and would not wait for the legacy stream)
  if (!stream && hipStreamSynchronize(nullptr) != hipSuccess) return LAM_ERR_HIP;
  return LAM_OK;
}

int lam_eigh_warm(const float *d_A, const float *d_prop, int n, int nmat, int warm_vectors, int k,
                  int steps, float tol, float *d_evals, float *d_evecs, float *d_ratio, void *stream) {
  if (!d_A || !d_prop || n < 1 || n > kMaxReceivers || nmat < 1 || warm_vectors < 2 ||
      warm_vectors > kMaxWarm || (warm_vectors & 1) || n < 2 * warm_vectors || k < 0 || k > warm_vectors ||
      steps < 0 || !d_evals || !d_evecs || !d_ratio)
    return LAM_ERR_ARG;
  const hipError_t e = launch_eigh_warm(d_A, d_prop, n, nmat, warm_vectors, k, steps, tol, d_evals, d_evecs,
                                        d_ratio, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? LAM_OK : LAM_ERR_HIP;
}

int lam_eigh_batched(const float *d_A, int n, int nmat, int whiten, int max_sweeps, float tol,
                     float *d_evals, float *d_evecs, int32_t *d_sweeps, void *stream) {
  if (!d_A || n < 1 || n > kMaxReceivers || nmat < 1 || !d_evals || !d_evecs) return LAM_ERR_ARG;
  const hipError_t e = launch_eigh_batched(d_A, n, nmat, whiten, max_sweeps, tol, d_evals, d_evecs,
                                           d_sweeps, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? LAM_OK : LAM_ERR_HIP;
}

}  // extern "C"
