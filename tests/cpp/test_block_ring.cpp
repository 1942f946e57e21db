// CPU test of spatial/block_ring.hpp (no GPU): a fake pipeline stands in for
// the HIP one.  It "copies" each block when execute_pipeline is called (as
// lam_execute's H2D copy would) and releases the buffer later from another
// thread (as the HIP callback does), checking at release time that the
// producer has not written the block in between.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "spatial/block_ring.hpp"

static int g_fail = 0;
#define CHECK(cond, ...)                                                   \
  do {                                                                     \
    if (!(cond)) {                                                         \
      ++g_fail;                                                            \
      std::printf("  CHECK failed %s:%d: %s ", __FILE__, __LINE__, #cond); \
      std::printf(__VA_ARGS__);                                            \
      std::printf("\n");                                                   \
    }                                                                      \
  } while (0)

struct FakeBlock : FinalPacketData {
  int samples[16] = {0};
  int16_t scales[4] = {1, 1, 1, 1};
  bool arrivals[4] = {true, true, true, true};
  void *get_samples_ptr() override { return samples; }
  size_t get_samples_elements_size() override { return sizeof(samples); }
  void *get_scales_ptr() override { return scales; }
  size_t get_scales_element_size() override { return sizeof(scales); }
  bool *get_arrivals_ptr() override { return arrivals; }
  size_t get_arrivals_size() override { return sizeof(arrivals); }
  void zero_missing_packets() override {}
  int get_num_missing_packets() override { return 0; }
};
struct FakeConfig {
  using PacketFinalDataType = FakeBlock;
};

class FakePipeline : public GPUPipeline {
 public:
  explicit FakePipeline(int fail_at = -1) : fail_at_(fail_at) {}
  ~FakePipeline() override {
    for (auto &t : releasers_) t.join();
  }
  void execute_pipeline(FinalPacketData *pd, const bool = false) override {
    auto *b = static_cast<FakeBlock *>(pd);
    const int n = calls_++;
    if (n == fail_at_) {
      // as lam_execute does when it fails after queuing the release: the
      // block comes back once, later, from the callback thread
      const size_t idx = pd->buffer_index;
      releasers_.emplace_back([this, idx] {
        std::this_thread::sleep_for(std::chrono::milliseconds(5));
        state_->release_buffer(idx);
      });
      throw std::runtime_error("injected pipeline failure");
    }
    const int tag = b->samples[0];
    {
      std::lock_guard<std::mutex> lock(mu_);
      executed_.push_back(tag);
      for (int i = 1; i < 16; ++i) CHECK(b->samples[i] == tag, "block %d torn at execute", tag);
    }
    const size_t idx = pd->buffer_index;
    releasers_.emplace_back([this, b, tag, idx] {
      std::this_thread::sleep_for(std::chrono::microseconds(300 + 97 * (tag % 5)));
      for (int i = 0; i < 16; ++i) CHECK(b->samples[i] == tag, "block %d overwritten before its release", tag);
      state_->release_buffer(idx);
    });
  }
  void dump_visibilities(const uint64_t = 0) override {}
  std::vector<int> executed() {
    std::lock_guard<std::mutex> lock(mu_);
    return executed_;
  }

 private:
  int fail_at_;
  std::atomic<int> calls_{0};
  std::mutex mu_;
  std::vector<int> executed_;
  std::vector<std::thread> releasers_;
};

static void order_and_reuse() {
  constexpr int kBlocks = 40;
  FakePipeline pipeline;
  {
    BlockRing<FakeConfig> ring(3);
    ring.start(pipeline);
    for (int b = 0; b < kBlocks; ++b) {
      const size_t i = ring.acquire();
      for (int k = 0; k < 16; ++k) ring.block(i).samples[k] = b;
      ring.block(i).buffer_index = 99;  // the ring restamps its own slot
      ring.submit(i, 10 * b, 10 * b + 9);
    }
    ring.stop();
    CHECK(ring.blocks_executed() == (size_t)kBlocks, "executed %zu", ring.blocks_executed());
  }
  const auto ex = pipeline.executed();
  CHECK(ex.size() == (size_t)kBlocks, "size %zu", ex.size());
  for (size_t k = 0; k < ex.size(); ++k) CHECK(ex[k] == (int)k, "order at %zu: %d", k, ex[k]);
}

static void failure_does_not_strand_the_producer() {
  FakePipeline pipeline(2);  // the third block throws
  BlockRing<FakeConfig> ring(2);
  ring.start(pipeline);
  int submitted = 0, refused = 0;
  for (int b = 0; b < 10; ++b) {
    try {
      const size_t i = ring.acquire();
      for (int k = 0; k < 16; ++k) ring.block(i).samples[k] = b;
      ring.submit(i, 0, 0);
      ++submitted;
    } catch (const std::runtime_error &) {
      ++refused;
    }
  }
  bool rethrown = false;
  try {
    ring.stop();
  } catch (const std::runtime_error &e) {
    rethrown = std::strstr(e.what(), "injected") != nullptr;
  }
  CHECK(rethrown, "stop() rethrows the pipeline's exception");
  CHECK(refused > 0 && submitted + refused == 10, "submitted %d refused %d", submitted, refused);
  // stop() returned after the failed block's release, and every buffer is
  // free exactly once (a buffer pushed by both the feeder and the release
  // would be handed out twice)
  CHECK(ring.blocks_released() == ring.blocks_executed(), "released %zu of %zu", ring.blocks_released(),
        ring.blocks_executed());
  size_t handed = 0;
  try {
    for (int k = 0; k < 8; ++k) {
      ring.acquire();
      ++handed;
    }
  } catch (const std::runtime_error &) {
  }
  CHECK(handed == ring.size(), "free buffers after the failure: %zu of %zu", handed, ring.size());
}

int main() {
  order_and_reuse();
  failure_does_not_strand_the_producer();
  std::printf("%d failures\n", g_fail);
  return g_fail ? 1 : 0;
}
