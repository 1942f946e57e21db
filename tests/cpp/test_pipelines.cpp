// C++ seam tests of LambdaCorrBeamOnlyGPUPipeline / LambdaAdaptiveBeamformedSpectraPipeline
// (cuda-spatial-filtering_amd/include/spatial/pipeline.hpp) on the GPU.
// They drive execute_pipeline with reassembled blocks the way the reference's
// SyntheticPipelineRun does (tests/support/pipeline_harness.hpp:171-260:
// packet k of the window in ring slot k+1, packet -1 in slot 0) and assert the
// reference's golden values (tests/test_corr_beam_pipeline.cu,
// tests/test_adaptive_pipeline.cu).
#include <cmath>
#include <complex>
#include <cstdio>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "spatial/block_ring.hpp"
#include "spatial/output.hpp"
#include "spatial/packet_formats.hpp"
#include "spatial/pipeline.hpp"

static int g_fail = 0, g_checks = 0;
#define CHECK(cond, ...)                                                   \
  do {                                                                     \
    ++g_checks;                                                            \
    if (!(cond)) {                                                         \
      ++g_fail;                                                            \
      std::printf("  CHECK failed %s:%d: %s ", __FILE__, __LINE__, #cond); \
      std::printf(__VA_ARGS__);                                            \
      std::printf("\n");                                                   \
    }                                                                      \
  } while (0)
#define TEST(name) static void name()
#define RUN(name)                                  \
  do {                                             \
    const int before = g_fail;                     \
    std::printf("[ RUN  ] %s\n", #name);           \
    std::fflush(stdout);                           \
    name();                                        \
    std::printf("[ %s ] %s\n", g_fail == before ? " OK " : "FAIL", #name); \
    std::fflush(stdout);                           \
  } while (0)

// test_configs.hpp:17-78 and test_adaptive_pipeline.cu:37-62
using Small = LambdaConfig<1, 1, 8, 4, 2, 4, 1, 1, 32, 32, 10000>;
using TwoChannel = LambdaConfig<2, 1, 8, 4, 2, 4, 1, 1, 32, 32, 10000>;
using TwoPacket = LambdaConfig<1, 1, 8, 4, 2, 4, 2, 1, 32, 32, 10000>;
using MultiFPGA = LambdaConfig<1, 3, 8, 6, 2, 2, 1, 1, 32, 32, 10000>;
using AutoDump = LambdaConfig<1, 1, 8, 4, 2, 4, 1, 1, 32, 32, 2>;
using Adaptive = LambdaConfig<1, 1, 64, 4, 2, 4, 1, 1, 32, 32, 1>;
using SmallGPU = LambdaConfig<1, 1, 8, 4, 2, 4, 1, 1, 32, 32, 10000, false, 2>;  // bandpass of 2-bin runs
using Adaptive2 = LambdaConfig<2, 1, 64, 4, 2, 4, 1, 1, 32, 32, 1>;

struct FakeProcessorState : ProcessorStateBase {
  std::mutex m;
  std::vector<size_t> released;
  void release_buffer(const size_t i) override {
    std::lock_guard<std::mutex> lock(m);
    released.push_back(i);
  }
};

using SampleFn = std::function<std::complex<int8_t>(size_t, size_t, int, int, int, int)>;
using ScaleFn = std::function<int16_t(size_t, size_t, int, int, int)>;

template <typename Cfg>
void fill_block(typename Cfg::PacketFinalDataType &pd, const SampleFn &sample, const ScaleFn &scale,
                const std::vector<std::vector<int>> &missing = {}) {
  for (size_t c = 0; c < Cfg::NR_CHANNELS; ++c)
    for (size_t slot = 0; slot < Cfg::NR_PACKETS_FOR_CORRELATION + 2; ++slot)
      for (size_t f = 0; f < Cfg::NR_FPGA_SOURCES; ++f) {
        const int pkt = static_cast<int>(slot) - 1;
        (*pd.arrivals)[c][slot][f] = true;
        for (size_t n = 0; n < Cfg::NR_RECEIVERS_PER_PACKET; ++n)
          for (size_t p = 0; p < Cfg::NR_POLARIZATIONS; ++p) {
            (*pd.scales)[c][slot][f * Cfg::NR_RECEIVERS_PER_PACKET + n][p] =
                scale(c, f, pkt, (int)n, (int)p);
            for (size_t t = 0; t < Cfg::NR_TIME_STEPS_PER_PACKET; ++t)
              (*pd.samples)[c][slot][f][t][n][p] = sample(c, f, pkt, (int)t, (int)n, (int)p);
          }
      }
  for (const auto &m : missing) (*pd.arrivals)[m[0]][m[1]][m[2]] = false;
  pd.zero_missing_packets();
  pd.start_seq_id = 1000;
  pd.end_seq_id = 1000 + Cfg::NR_PACKETS_FOR_CORRELATION * Cfg::NR_TIME_STEPS_PER_PACKET;
  pd.buffer_index = 7;
}

template <typename Cfg>
BeamWeightsT<Cfg> unity_weights(float v = 1.0f) {
  BeamWeightsT<Cfg> w{};
  for (auto &a : w.weights)
    for (auto &b : a)
      for (auto &c : b)
        for (auto &d : c) d = spatial::complex_half(v, 0.0f);
  return w;
}

static size_t bl(size_t i, size_t j) { return j * (j + 1) / 2 + i; }

static std::complex<int8_t> const_sample(size_t, size_t, int, int, int, int) { return {2, -2}; }
static int16_t unit_scale(size_t, size_t, int, int, int) { return 1; }

template <typename Cfg>
struct CorrBeamRun {
  std::shared_ptr<SingleHostMemoryOutput<Cfg>> output = std::make_shared<SingleHostMemoryOutput<Cfg>>();
  BeamWeightsT<Cfg> weights;
  FakeProcessorState state;
  std::unique_ptr<LambdaCorrBeamOnlyGPUPipeline<Cfg>> pipeline;
  explicit CorrBeamRun(BeamWeightsT<Cfg> w) : weights(w) {
    pipeline = std::make_unique<LambdaCorrBeamOnlyGPUPipeline<Cfg>>(
        (int)Cfg::NR_PACKETS_FOR_CORRELATION, &weights);
    pipeline->set_state(&state);
    pipeline->set_output(output);
  }
  void run(const SampleFn &s, const ScaleFn &sc, bool dump = true,
           const std::vector<std::vector<int>> &missing = {}) {
    typename Cfg::PacketFinalDataType pd;
    fill_block<Cfg>(pd, s, sc, missing);
    pipeline->execute_pipeline(&pd);
    if (dump) pipeline->dump_visibilities();
    pipeline->synchronize();
  }
};

TEST(CorrBeamExactValues) {
  CorrBeamRun<Small> r(unity_weights<Small>());
  r.run(const_sample, unit_scale);
  auto &beam = *r.output->beam_data;
  auto &vis = *r.output->visibilities;
  for (size_t t = 0; t < 8; ++t)
    for (size_t p = 0; p < 2; ++p) {
      CHECK(spatial::half_to_float(beam[0][p][0][t][0]) == 8.0f, "t=%zu", t);
      CHECK(spatial::half_to_float(beam[0][p][0][t][1]) == -8.0f, "t=%zu", t);
    }
  for (size_t b = 0; b < Small::NR_BASELINES_UNPADDED; ++b)
    for (size_t p = 0; p < 2; ++p)
      for (size_t q = 0; q < 2; ++q) {
        CHECK(vis[0][b][p][q][0] == 64.0f, "bl=%zu", b);
        CHECK(vis[0][b][p][q][1] == 0.0f, "bl=%zu", b);
      }
  CHECK(r.state.released.size() == 1 && r.state.released[0] == 7, "release count %zu",
        r.state.released.size());
  CHECK(r.output->beam_blocks_done == 1 && r.output->vis_blocks_done == 1, "callbacks");
  CHECK(r.output->last_total_packets == 1 && r.output->last_missing_packets == 0, "packets");
}

TEST(CorrBeamZeroInputAndWeights) {
  {
    CorrBeamRun<Small> r(unity_weights<Small>());
    r.run([](size_t, size_t, int, int, int, int) { return std::complex<int8_t>{0, 0}; }, unit_scale);
    for (auto &a : *r.output->visibilities)
      for (auto &b : a)
        for (auto &c : b)
          for (auto &d : c) CHECK(d[0] == 0.0f && d[1] == 0.0f, "vis");
  }
  {
    CorrBeamRun<Small> r(BeamWeightsT<Small>{});
    r.run(const_sample, unit_scale);
    for (size_t t = 0; t < 8; ++t) CHECK(spatial::half_to_float((*r.output->beam_data)[0][0][0][t][0]) == 0.0f, "beam");
    for (size_t rx = 0; rx < 4; ++rx) CHECK((*r.output->visibilities)[0][bl(rx, rx)][0][0][0] > 0.0f, "auto");
  }
  {
    CorrBeamRun<Small> r(unity_weights<Small>(2.0f));
    r.run(const_sample, unit_scale);
    CHECK(spatial::half_to_float((*r.output->beam_data)[0][1][0][3][0]) == 16.0f, "beam 16");
    CHECK(spatial::half_to_float((*r.output->beam_data)[0][1][0][3][1]) == -16.0f, "beam -16");
  }
}

TEST(CorrBeamTwoChannelsScalesPackets) {
  {
    CorrBeamRun<TwoChannel> r(unity_weights<TwoChannel>());
    r.run([](size_t c, size_t, int, int, int, int) {
      return c == 0 ? std::complex<int8_t>{2, -2} : std::complex<int8_t>{3, 1};
    }, unit_scale);
    CHECK(spatial::half_to_float((*r.output->beam_data)[1][0][0][0][0]) == 12.0f, "ch1 re");
    CHECK(spatial::half_to_float((*r.output->beam_data)[1][0][0][0][1]) == 4.0f, "ch1 im");
    CHECK((*r.output->visibilities)[1][bl(2, 2)][1][1][0] == 80.0f, "auto 80");
  }
  {
    CorrBeamRun<Small> r(unity_weights<Small>());
    r.run([](size_t, size_t, int, int, int, int) { return std::complex<int8_t>{1, 0}; },
          [](size_t, size_t, int, int rx, int p) { return (int16_t)((rx + 1) * (p + 1)); });
    CHECK(spatial::half_to_float((*r.output->beam_data)[0][0][0][5][0]) == 10.0f, "beam 10");
    CHECK(spatial::half_to_float((*r.output->beam_data)[0][1][0][5][0]) == 20.0f, "beam 20");
    for (int rx = 0; rx < 4; ++rx)
      for (int p = 0; p < 2; ++p) {
        const float s = (float)((rx + 1) * (p + 1));
        CHECK((*r.output->visibilities)[0][bl(rx, rx)][p][p][0] == 8.0f * s * s, "N s^2");
      }
  }
  {
    CorrBeamRun<TwoPacket> r(unity_weights<TwoPacket>());
    r.run(const_sample, unit_scale);
    CHECK((*r.output->visibilities)[0][bl(3, 3)][0][0][0] == 128.0f, "two packets");
  }
}

TEST(CorrBeamMultiFPGAAndMissingPackets) {
  // 3 FPGAs x 2 receivers; packet -1 of FPGA 1 missing -> its scales are zero
  CorrBeamRun<MultiFPGA> r(unity_weights<MultiFPGA>());
  r.run([](size_t, size_t f, int, int, int n, int) {
    return std::complex<int8_t>{(int8_t)(1 + f), (int8_t)n};
  }, unit_scale, true, {{0, 0, 1}});
  auto &vis = *r.output->visibilities;
  // receiver 0 (FPGA 0, n=0): (1,0) over 8 samples -> 8
  CHECK(vis[0][bl(0, 0)][0][0][0] == 8.0f, "fpga0 auto %f", vis[0][bl(0, 0)][0][0][0]);
  // receiver 2 (FPGA 1) blanked by the missing packet -> 0
  CHECK(vis[0][bl(2, 2)][0][0][0] == 0.0f, "blanked auto %f", vis[0][bl(2, 2)][0][0][0]);
  // receiver 5 (FPGA 2, n=1): (3,1) -> 8*10
  CHECK(vis[0][bl(5, 5)][1][1][0] == 80.0f, "fpga2 auto");
  // cross 0x5: sum (1)(3-1i) = 8*(3-1i)
  CHECK(vis[0][bl(0, 5)][0][0][0] == 24.0f && vis[0][bl(0, 5)][0][0][1] == -8.0f, "cross");
}

TEST(CorrBeamAutomaticDumpAndStateChecks) {
  CorrBeamRun<AutoDump> r(unity_weights<AutoDump>());
  r.run(const_sample, unit_scale, false);
  CHECK(r.output->vis_blocks_done == 0, "no dump after 1 block");
  r.run(const_sample, unit_scale, false);
  CHECK(r.output->vis_blocks_done == 1, "dump after 2 blocks (%d)", r.output->vis_blocks_done.load());
  CHECK((*r.output->visibilities)[0][bl(1, 1)][0][0][0] == 128.0f, "2 blocks accumulated");
  CHECK(r.output->last_total_packets == 2, "total packets %d", r.output->last_total_packets);
  BeamWeightsT<Small> w = unity_weights<Small>();
  LambdaCorrBeamOnlyGPUPipeline<Small> p(1, &w);
  Small::PacketFinalDataType pd;
  bool threw = false;
  try {
    p.execute_pipeline(&pd);
  } catch (const std::logic_error &) {
    threw = true;
  }
  CHECK(threw, "logic_error without state");
}

template <typename Cfg>
struct AdaptiveRun {
  std::shared_ptr<SingleHostMemoryOutput<Cfg, 2 * Cfg::NR_BEAMS>> output =
      std::make_shared<SingleHostMemoryOutput<Cfg, 2 * Cfg::NR_BEAMS>>();
  BeamWeightsT<Cfg> weights = unity_weights<Cfg>();
  FakeProcessorState state;
  std::unique_ptr<LambdaAdaptiveBeamformedSpectraPipeline<Cfg>> pipeline;
  AdaptiveRun(bool shrink, int K, bool detect = false, float delta = 3.0f) {
    std::unordered_map<int, int> eig_map;
    for (int c = 0; c < (int)Cfg::NR_CHANNELS; ++c) eig_map[c] = K;
    pipeline = std::make_unique<LambdaAdaptiveBeamformedSpectraPipeline<Cfg>>(
        3, &weights, eig_map, 0, BeamSteering<Cfg>{}, shrink, detect, delta);
    pipeline->set_state(&state);
    pipeline->set_output(output);
  }
  void run(const SampleFn &s) {
    typename Cfg::PacketFinalDataType pd;
    fill_block<Cfg>(pd, s, unit_scale);
    pipeline->execute_pipeline(&pd);
    pipeline->synchronize();
  }
  double mitigated_power(int pol) {
    double t = 0;
    for (size_t c = 0; c < Cfg::NR_CHANNELS; ++c)
      for (size_t i = 0; i < 64; ++i) {
        const float re = spatial::half_to_float((*output->beam_data)[c][pol][1][i][0]);
        const float im = spatial::half_to_float((*output->beam_data)[c][pol][1][i][1]);
        t += re * re + im * im;
      }
    return t;
  }
};

TEST(AdaptiveNullRankOne) {
  AdaptiveRun<Adaptive> r(false, 1);
  r.run(const_sample);
  auto &beam = *r.output->beam_data;
  for (int p = 0; p < 2; ++p)
    for (int t = 0; t < 64; ++t) {
      CHECK(spatial::half_to_float(beam[0][p][0][t][0]) == 8.0f, "orig re");
      CHECK(std::fabs(spatial::half_to_float(beam[0][p][1][t][0])) <= 0.25f, "mitigated re");
      CHECK(std::fabs(spatial::half_to_float(beam[0][p][1][t][1])) <= 0.25f, "mitigated im");
    }
  CHECK(r.mitigated_power(0) + r.mitigated_power(1) < 1.0, "power");
  // the adaptive pipeline's [C][P][R] eigenvalues at the start of the landing area
  auto &ev = *reinterpret_cast<float(*)[Adaptive::NR_CHANNELS][2][Adaptive::NR_RECEIVERS]>(r.output->eigenvalues);
  for (int p = 0; p < 2; ++p)
    for (int k = 0; k < 4; ++k) {
      CHECK(ev[0][p][k] >= -1e-3f, "non-negative");
      if (k) CHECK(ev[0][p][k] >= ev[0][p][k - 1] - 1e-3f, "ascending");
    }
  CHECK((*r.output->nulled_eigenmode_counts)[0][0] == 1, "count");
  CHECK(r.output->eig_blocks_done == 1 && r.output->beam_blocks_done == 1, "callbacks");
  // spectra: a constant beam is all DC, which lands on output bin 27 (adaptive 816-833)
  CHECK(r.output->fft_blocks_done == 1, "fft callback");
  auto &fft = *r.output->fft_output;
  const float dc = 64.0f * std::sqrt(128.0f);
  for (int p = 0; p < 2; ++p)
    for (int g = 0; g < 54; ++g) {
      CHECK(std::isfinite(fft[0][p][0][g]) && std::isfinite(fft[0][p][1][g]), "finite");
      if (g == 27)
        CHECK(std::fabs(fft[0][p][0][g] - dc) < 1e-3f * dc, "dc %f", fft[0][p][0][g]);
      else
        CHECK(fft[0][p][0][g] < 1e-3f * dc, "bin %d = %f", g, fft[0][p][0][g]);
    }
  CHECK(r.state.released.size() == 1, "released once");
}

TEST(AdaptiveK0TaggedLayout) {
  auto tagged = [](size_t c, size_t, int, int t, int rx, int p) {
    if (rx != 0) return std::complex<int8_t>{0, 0};
    return std::complex<int8_t>{(int8_t)(t + 1), (int8_t)(-(1 + 10 * p + 20 * (int)c))};
  };
  for (int shrink = 0; shrink < 2; ++shrink) {
    AdaptiveRun<Adaptive2> r(shrink, 0);
    r.run(tagged);
    auto &beam = *r.output->beam_data;
    for (int c = 0; c < 2; ++c)
      for (int p = 0; p < 2; ++p)
        for (int b = 0; b < 2; ++b)
          for (int t = 0; t < 64; ++t) {
            CHECK(spatial::half_to_float(beam[c][p][b][t][0]) == (float)(t + 1), "re c=%d p=%d b=%d t=%d", c, p, b, t);
            CHECK(spatial::half_to_float(beam[c][p][b][t][1]) == (float)(-(1 + 10 * p + 20 * c)), "im");
          }
  }
}

TEST(AdaptiveShrinkAndDetection) {
  auto rank2 = [](size_t, size_t, int, int t, int rx, int) {
    if (t < 32) return std::complex<int8_t>{2, -2};
    return rx == 0 ? std::complex<int8_t>{4, 0} : std::complex<int8_t>{0, 0};
  };
  AdaptiveRun<Adaptive> rn(false, 1), rs(true, 1);
  rn.run(rank2);
  rs.run(rank2);
  CHECK(rs.mitigated_power(0) > rn.mitigated_power(0) + 1e-6, "shrink > null");
  auto multimode = [](size_t, size_t, int, int t, int rx, int p) {
    if (p == 1) return std::complex<int8_t>{0, 0};
    const int a = ((t / 8) % 2 == 0) ? 20 : -20, b = ((t / 4) % 2 == 0) ? 18 : -18;
    return std::complex<int8_t>{(int8_t)(a + (rx % 2 == 0 ? 1 : -1) * b), 0};
  };
  AdaptiveRun<Adaptive> rd(false, 0, true, 0.5f);
  rd.run(multimode);
  CHECK((*rd.output->nulled_eigenmode_counts)[0][0] == 2, "counts[0]=%d", (*rd.output->nulled_eigenmode_counts)[0][0]);
  CHECK((*rd.output->nulled_eigenmode_counts)[0][1] == 0, "counts[1]");
  CHECK(rd.mitigated_power(0) < 1.0, "interfered pol nulled");
  CHECK(rd.mitigated_power(1) == 0.0, "clean pol zero");
}

TEST(AdaptivePermutationSizeError) {
  AdaptiveRun<Adaptive> r(false, 1);
  bool threw = false;
  try {
    r.pipeline->set_stream_permutation({0, 1}, {0, 1});
  } catch (const std::runtime_error &) {
    threw = true;
  }
  CHECK(threw, "runtime_error on size mismatch");
  // identity permutation is accepted and leaves results unchanged
  std::vector<int> recv, pol;
  for (int i = 0; i < 4; ++i)
    for (int p = 0; p < 2; ++p) {
      recv.push_back(i);
      pol.push_back(p);
    }
  r.pipeline->set_stream_permutation(recv, pol);
  r.run(const_sample);
  CHECK(spatial::half_to_float((*r.output->beam_data)[0][0][0][0][0]) == 8.0f, "identity perm");
}

// LambdaGPUPipeline, tests/test_pipeline.cu:956-1052 (EigenvalueBasic): every
// sample (2, -2), 4 receivers, 8 samples: each pol-pair covariance is 64 in
// every entry, so its eigenvalues are 0, 0, 0, 256 and the principal
// eigenvector is 0.5 in every component.  The other three eigenvectors span a
// degenerate eigenspace (any basis of it is a valid answer), so only their
// orthogonality to the principal one is checked.  Beams and visibilities as
// in test_pipeline.cu's Ex1.
TEST(LambdaGPUEigenvalueBasic) {
  auto output = std::make_shared<SingleHostMemoryOutput<SmallGPU>>();
  BeamWeightsT<SmallGPU> weights = unity_weights<SmallGPU>();
  LambdaGPUPipeline<SmallGPU> pipeline(1, &weights);
  FakeProcessorState state;
  pipeline.set_state(&state);
  pipeline.set_output(output);
  int delays[1] = {0};
  pipeline.set_subpacket_delays(delays);
  typename SmallGPU::PacketFinalDataType pd;
  fill_block<SmallGPU>(pd, const_sample, unit_scale);
  pipeline.execute_pipeline(&pd);
  pipeline.dump_visibilities();
  pipeline.synchronize();
  auto &ev = *output->eigenvalues;
  auto &vec = *output->eigenvectors;
  for (int p = 0; p < 2; ++p)
    for (int q = 0; q < 2; ++q) {
      for (int n = 0; n < 4; ++n)
        CHECK(std::fabs(ev[0][p][q][n] - (n == 3 ? 256.0f : 0.0f)) <= 1e-4f, "eig p=%d q=%d n=%d %f", p, q,
              n, ev[0][p][q][n]);
      for (int l = 0; l < 4; ++l) {
        CHECK(std::fabs(vec[0][p][q][3][l].real() - 0.5f) <= 1e-4f, "principal %f", vec[0][p][q][3][l].real());
        CHECK(std::fabs(vec[0][p][q][3][l].imag()) <= 1e-4f, "principal imag");
      }
      for (int n = 0; n < 3; ++n) {
        std::complex<float> dot = 0.0f;
        for (int l = 0; l < 4; ++l) dot += std::conj(vec[0][p][q][3][l]) * vec[0][p][q][n][l];
        CHECK(std::abs(dot) <= 1e-4f, "orthogonal");
      }
    }
  CHECK(output->eig_blocks_done == 1, "eigen callback");
  // bandpass: the constant beam 8 - 8i is all DC, |X_0|^2 = (8 * |8 - 8i|)^2 = 8192,
  // averaged with bin 1 (0) into output bin 0
  CHECK(output->fft_blocks_done == 1, "fft callback");
  for (int p = 0; p < 2; ++p)
    for (int m = 0; m < 4; ++m)
      CHECK(std::fabs(output->bandpass()[0][p][0][m] - (m == 0 ? 4096.0f : 0.0f)) <= 1e-3f,
            "bandpass m=%d %f", m, output->bandpass()[0][p][0][m]);
  for (size_t t = 0; t < 8; ++t) CHECK(spatial::half_to_float((*output->beam_data)[0][0][0][t][0]) == 8.0f, "beam");
  CHECK((*output->visibilities)[0][0][0][0][0] == 64.0f, "vis");
}

// LambdaProjectionPipeline: every sample (2, -2) makes each pol-pair
// covariance 64 everywhere, its top eigenvector 0.5 * ones, so each block's
// projector (K = 1) is 0.25 everywhere; after NR_RUNS_TO_AVERAGE = 2 blocks the
// average is dumped: eigenvalues 0, 0, 0, 1 and principal eigenvector 0.5.
TEST(ProjectionAverageBasic) {
  auto output = std::make_shared<SingleHostMemoryOutput<Small>>();
  LambdaProjectionPipeline<Small, 1, 2> pipeline(2);
  FakeProcessorState state;
  pipeline.set_state(&state);
  pipeline.set_output(output);
  for (int b = 0; b < 2; ++b) {
    typename Small::PacketFinalDataType pd;
    fill_block<Small>(pd, const_sample, unit_scale);
    pipeline.execute_pipeline(&pd);
    pipeline.synchronize();
  }
  CHECK(output->eig_blocks_done == 1, "one projection dump after two blocks (%d)", output->eig_blocks_done.load());
  auto &ev = *output->eigenvalues;
  auto &vec = *output->eigenvectors;
  for (int p = 0; p < 2; ++p)
    for (int q = 0; q < 2; ++q) {
      for (int n = 0; n < 4; ++n)
        CHECK(std::fabs(ev[0][p][q][n] - (n == 3 ? 1.0f : 0.0f)) <= 1e-4f, "eig %d %f", n, ev[0][p][q][n]);
      for (int l = 0; l < 4; ++l)
        CHECK(std::fabs(vec[0][p][q][3][l].real() - 0.5f) <= 1e-4f && std::fabs(vec[0][p][q][3][l].imag()) <= 1e-4f,
              "principal");
    }
  CHECK(state.released.size() == 2, "released twice");
}

// Overlapped ingest (ProcessorState -> pinned buffers -> async H2D,
// spatial.hpp:718-770, 1182-1299): 12 blocks through 3 pinned buffers, the
// producer refilling each buffer the moment the pipeline releases it.  A
// release issued before that block's H2D copy finished would let the next
// block's data into the copy and break the exact visibility sum.
TEST(BlockRingOverlappedIngest) {
  constexpr int kBlocks = 12;
  auto output = std::make_shared<SingleHostMemoryOutput<Small>>();
  BeamWeightsT<Small> weights = unity_weights<Small>();
  LambdaCorrBeamOnlyGPUPipeline<Small> pipeline(3, &weights);
  pipeline.set_output(output);
  BlockRing<Small> ring(3);
  ring.start(pipeline);
  for (int b = 0; b < kBlocks; ++b) {
    const size_t i = ring.acquire();
    fill_block<Small>(ring.block(i),
                      [b](size_t, size_t, int, int, int, int) { return std::complex<int8_t>{(int8_t)(b + 1), 0}; },
                      unit_scale);
    ring.submit(i, 100 * (uint64_t)b, 100 * (uint64_t)b + 99);
  }
  ring.stop();
  pipeline.dump_visibilities();
  pipeline.synchronize();
  CHECK(ring.blocks_executed() == (size_t)kBlocks, "executed %zu", ring.blocks_executed());
  CHECK(ring.blocks_released() == (size_t)kBlocks, "released %zu", ring.blocks_released());
  CHECK(ring.missing_packets() == 0, "missing");
  float expect = 0.0f;
  for (int b = 0; b < kBlocks; ++b) expect += 8.0f * (float)((b + 1) * (b + 1));
  auto &vis = *output->visibilities;
  for (size_t bl_ = 0; bl_ < Small::NR_BASELINES_UNPADDED; ++bl_)
    for (size_t p = 0; p < 2; ++p)
      for (size_t q = 0; q < 2; ++q)
        CHECK(vis[0][bl_][p][q][0] == expect && vis[0][bl_][p][q][1] == 0.0f, "vis %f", vis[0][bl_][p][q][0]);
  // the landing area holds one whole block's beams (4 receivers x (b + 1))
  auto &beam = *output->beam_data;
  const float b0 = spatial::half_to_float(beam[0][0][0][0][0]);
  CHECK(b0 >= 4.0f && b0 <= 4.0f * kBlocks && std::fmod(b0, 4.0f) == 0.0f, "beam %f", b0);
  for (size_t p = 0; p < 2; ++p)
    for (size_t t = 0; t < 8; ++t) CHECK(spatial::half_to_float(beam[0][p][0][t][0]) == b0, "whole block");
  CHECK(output->beam_blocks_done == kBlocks, "beam callbacks %d", output->beam_blocks_done.load());
}

// The reference harness's factories, written exactly as
// tests/support/pipeline_harness.hpp:73-96 and 145-165 write them: inert
// BeamSteering from its full constructor, then the pipeline constructors.
namespace pipeline_factories {
template <typename Config>
std::unique_ptr<LambdaGPUPipeline<Config>> make_gpu_pipeline(int num_buffers, BeamWeightsT<Config> *weights,
                                                             int min_freq_channel = 0) {
  BeamSteering<Config> beam_steering(/*targets=*/{}, /*antenna_positions=*/{},
                                     /*antenna_mapping=*/{}, FrequencyPlan{}, min_freq_channel, ArrayLocation{},
                                     /*update_interval_seconds=*/1.0, num_buffers);
  return std::make_unique<LambdaGPUPipeline<Config>>(num_buffers, weights, std::move(beam_steering));
}
template <typename Config>
std::unique_ptr<LambdaCorrBeamOnlyGPUPipeline<Config>> make_corr_beam_only_pipeline(int num_buffers,
                                                                                     BeamWeightsT<Config> *weights,
                                                                                     int min_freq_channel = 0) {
  BeamSteering<Config> beam_steering(/*targets=*/{}, /*antenna_positions=*/{},
                                     /*antenna_mapping=*/{}, FrequencyPlan{}, min_freq_channel, ArrayLocation{},
                                     /*update_interval_seconds=*/1.0, num_buffers);
  return std::make_unique<LambdaCorrBeamOnlyGPUPipeline<Config>>(num_buffers, weights, std::move(beam_steering));
}
template <typename Config>
std::unique_ptr<LambdaAdaptiveBeamformedSpectraPipeline<Config>> make_adaptive_beamformed_spectra_pipeline(
    int num_buffers, BeamWeightsT<Config> *weights, std::unordered_map<int, int> nr_signal_eigenvectors = {},
    bool shrink_eigenvalues = false, int min_freq_channel = 0, bool detect_signal_eigenmodes = false,
    float detection_threshold_delta = 3.0f, double eigenmode_stats_interval_seconds = 10.0) {
  if (nr_signal_eigenvectors.empty()) nr_signal_eigenvectors[min_freq_channel] = 1;
  BeamSteering<Config> beam_steering(/*targets=*/{}, /*antenna_positions=*/{},
                                     /*antenna_mapping=*/{}, FrequencyPlan{}, min_freq_channel, ArrayLocation{},
                                     /*update_interval_seconds=*/1.0, num_buffers);
  return std::make_unique<LambdaAdaptiveBeamformedSpectraPipeline<Config>>(
      num_buffers, weights, nr_signal_eigenvectors, min_freq_channel, std::move(beam_steering),
      shrink_eigenvalues, detect_signal_eigenmodes, detection_threshold_delta, eigenmode_stats_interval_seconds);
}
}  // namespace pipeline_factories

TEST(HarnessFactoriesAndRunEvents) {
  FakeProcessorState state;
  {  // corr-beam through the harness factory: beams (8, -8), vis 64 (test_corr_beam_pipeline.cu:107-140)
    BeamWeightsT<Small> w = unity_weights<Small>();
    auto p = pipeline_factories::make_corr_beam_only_pipeline<Small>(2, &w);
    auto out = std::make_shared<SingleHostMemoryOutput<Small>>();
    p->set_state(&state);
    p->set_output(out);
    for (int k = 0; k < 3; ++k) {
      typename Small::PacketFinalDataType pd;
      fill_block<Small>(pd, const_sample, unit_scale);
      p->execute_pipeline(&pd);
    }
    p->dump_visibilities();
    p->synchronize();
    CHECK(spatial::half_to_float((*out->beam_data)[0][1][0][5][0]) == 8.0f, "factory corr-beam beam");
    CHECK((*out->visibilities)[0][bl(2, 3)][0][1][0] == 3 * 64.0f, "factory corr-beam vis");
    // the benchmark ring: hipEventElapsedTime(start_run[i], stop_run[i]) as apps/bench_gpu.cu:226-232 reads it
    CHECK(p->benchmark_runs_done == 3, "runs %zu", p->benchmark_runs_done);
    for (size_t i = 0; i < p->benchmark_runs_done; ++i) {
      float ms = -1.0f;
      const hipError_t e = hipEventElapsedTime(&ms, p->start_run[i], p->stop_run[i]);
      CHECK(e == hipSuccess && ms >= 0.0f && ms < 1000.0f, "run %zu: %s %f ms", i, hipGetErrorString(e), ms);
    }
  }
  {  // adaptive through the harness factory with the default K map {min_freq_channel: 1}
    BeamWeightsT<Adaptive> w = unity_weights<Adaptive>();
    auto p = pipeline_factories::make_adaptive_beamformed_spectra_pipeline<Adaptive>(2, &w);
    auto out = std::make_shared<SingleHostMemoryOutput<Adaptive, 2 * Adaptive::NR_BEAMS>>();
    p->set_state(&state);
    p->set_output(out);
    typename Adaptive::PacketFinalDataType pd;
    fill_block<Adaptive>(pd, const_sample, unit_scale);
    p->execute_pipeline(&pd);
    p->synchronize();
    CHECK(spatial::half_to_float((*out->beam_data)[0][0][0][7][0]) == 8.0f, "factory adaptive original beam");
    double pw = 0.0;
    for (size_t t = 0; t < 64; ++t) {
      const float re = spatial::half_to_float((*out->beam_data)[0][0][1][t][0]);
      const float im = spatial::half_to_float((*out->beam_data)[0][0][1][t][1]);
      pw += re * re + im * im;
    }
    CHECK(pw < 1.0, "factory adaptive K=1 nulls the rank-one signal (%f)", pw);
  }
  {  // LambdaGPU through the harness factory: principal eigenvalue 256 (test_pipeline.cu:956-1052)
    BeamWeightsT<SmallGPU> w = unity_weights<SmallGPU>();
    auto p = pipeline_factories::make_gpu_pipeline<SmallGPU>(1, &w);
    auto out = std::make_shared<SingleHostMemoryOutput<SmallGPU>>();
    p->set_state(&state);
    p->set_output(out);
    typename SmallGPU::PacketFinalDataType pd;
    fill_block<SmallGPU>(pd, const_sample, unit_scale);
    p->execute_pipeline(&pd);
    p->synchronize();
    CHECK(std::fabs((*out->eigenvalues)[0][0][0][3] - 256.0f) <= 1e-4f, "factory LambdaGPU eigenvalue");
    float ms = -1.0f;
    CHECK(hipEventElapsedTime(&ms, p->start_run[0], p->stop_run[0]) == hipSuccess && ms >= 0.0f, "LambdaGPU run event");
  }
  {  // steering with targets needs the pointing geometry: refused loudly
    bool threw = false;
    try {
      BeamSteering<Small> s({BeamTarget{"radec", 83.6, 22.0}}, {}, {}, FrequencyPlan{}, 0, ArrayLocation{}, 1.0, 1);
    } catch (const std::runtime_error &) {
      threw = true;
    }
    CHECK(threw, "BeamSteering with targets throws");
    BeamSteering<Small> inert({}, {}, {}, FrequencyPlan{}, 0, ArrayLocation{}, 1.0, 1);
    CHECK(!inert.active() && !inert.maybe_refresh(), "inert steering");
  }
}

// A block whose scaled samples leave the exact int8 range (|x s| = 300) must
// not corrupt a feeder loop silently: the next execute_pipeline after it
// completes throws std::runtime_error, and every block is released once.
TEST(RangeErrorSurfacesInFeederLoop) {
  BeamWeightsT<Small> w = unity_weights<Small>();
  auto p = pipeline_factories::make_corr_beam_only_pipeline<Small>(2, &w);
  FakeProcessorState state;
  p->set_state(&state);
  typename Small::PacketFinalDataType bad, good;
  fill_block<Small>(bad, [](size_t, size_t, int, int, int, int) { return std::complex<int8_t>{100, 0}; },
                    [](size_t, size_t, int, int, int) { return (int16_t)3; });
  fill_block<Small>(good, const_sample, unit_scale);
  p->execute_pipeline(&bad);
  std::this_thread::sleep_for(std::chrono::milliseconds(300));
  bool threw = false;
  try {
    p->execute_pipeline(&good);
  } catch (const std::runtime_error &e) {
    threw = std::string(e.what()).find("int8 range") != std::string::npos;
  }
  CHECK(threw, "the block after a range error throws");
  p->execute_pipeline(&good);  // reported once; the stream carries on
  p->synchronize();
  CHECK(state.released.size() == 3, "released %zu", state.released.size());
}

int main() {
  RUN(CorrBeamExactValues);
  RUN(CorrBeamZeroInputAndWeights);
  RUN(CorrBeamTwoChannelsScalesPackets);
  RUN(CorrBeamMultiFPGAAndMissingPackets);
  RUN(CorrBeamAutomaticDumpAndStateChecks);
  RUN(AdaptiveNullRankOne);
  RUN(AdaptiveK0TaggedLayout);
  RUN(AdaptiveShrinkAndDetection);
  RUN(AdaptivePermutationSizeError);
  RUN(LambdaGPUEigenvalueBasic);
  RUN(ProjectionAverageBasic);
  RUN(BlockRingOverlappedIngest);
  RUN(HarnessFactoriesAndRunEvents);
  RUN(RangeErrorSurfacesInFeederLoop);
  std::printf("%d checks, %d failures\n", g_checks, g_fail);
  return g_fail == 0 ? 0 : 1;
}
