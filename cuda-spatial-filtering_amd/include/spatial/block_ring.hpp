// Pinned block ring between a reassembly engine and a GPUPipeline: the buffer
// life cycle of the reference's ProcessorState without its packet capture.
//
//   producer (reassembly)           feeder thread              HIP callback thread
//   i = acquire()   <-- free --------------------------------- release_buffer(i)
//   fill block(i)                                               (after the H2D copy)
//   submit(i, seq)  -- ready -->    execute_pipeline(block(i))
//
// acquire/release_buffer follow advance_to_next_buffer / release_buffer
// (include/spatial/spatial.hpp:718-770: a buffer is writable again only once
// the pipeline has copied it to the device, and its arrivals are cleared
// there); submit follows handle_buffer_completion (1182-1240: blank missing
// packets through their scales, stamp the block's sequence window, queue
// it); the feeder thread is pipeline_feeder (1275-1299).  Blocks are
// executed in submission order; with num_buffers >= 2 the producer fills
// one buffer while the pipeline's H2D copy and kernels of the others run.
// The blocks are pinned (lam_host_alloc), so the H2D copies are DMA at full
// link speed and overlap the previous blocks' kernels on their own streams.
// The ring relies on the pipeline releasing every block handed to
// execute_pipeline exactly once, also when execute_pipeline throws (the
// pipelines in spatial/pipeline.hpp do): a failed block comes back through
// release_buffer, never twice.
#pragma once
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "spatial/packet_formats.hpp"
#include "spatial/pipeline_base.hpp"

template <typename T>
class BlockRing : public ProcessorStateBase {
 public:
  using Block = typename T::PacketFinalDataType;

  explicit BlockRing(size_t num_buffers) {
    if (num_buffers < 1) throw std::invalid_argument("BlockRing needs at least one buffer");
    for (size_t i = 0; i < num_buffers; ++i) {
      blocks_.push_back(std::make_unique<Block>());
      blocks_.back()->buffer_index = i;
      free_.push_back(i);
    }
  }
  ~BlockRing() override {
    try {
      stop();
    } catch (...) {
    }
  }
  BlockRing(const BlockRing &) = delete;
  BlockRing &operator=(const BlockRing &) = delete;

  size_t size() const { return blocks_.size(); }
  Block &block(size_t i) { return *blocks_.at(i); }

  // Producer: wait until a buffer has been released by the pipeline and
  // return its index; the caller owns it until submit().
  // Throws once the feeder has stopped on a pipeline error and no buffer is
  // free (nothing would ever release one).
  size_t acquire() {
    std::unique_lock<std::mutex> lock(mu_);
    cv_free_.wait(lock, [this] { return !free_.empty() || failed_; });
    if (free_.empty()) throw std::runtime_error("BlockRing: the pipeline failed; no buffer will be released");
    const size_t i = free_.front();
    free_.pop_front();
    return i;
  }

  // Producer: hand a filled buffer to the feeder.
  void submit(size_t i, uint64_t start_seq, uint64_t end_seq) {
    Block &b = block(i);
    b.zero_missing_packets();
    const int missing = b.get_num_missing_packets();
    b.start_seq_id = start_seq;
    b.end_seq_id = end_seq;
    b.buffer_index = i;  // the ring's slot, whatever the producer wrote there
    {
      std::lock_guard<std::mutex> lock(mu_);
      if (stopping_) throw std::logic_error("BlockRing::submit after stop");
      if (failed_) {  // nothing executes any more: the buffer goes straight back
        free_.push_back(i);
        cv_free_.notify_one();
        throw std::runtime_error("BlockRing: the pipeline failed; block not executed");
      }
      missing_packets_ += missing;
      ready_.push_back(i);
    }
    cv_ready_.notify_one();
  }

  // Start the feeder thread on `pipeline` (which reports releases back here).
  void start(GPUPipeline &pipeline) {
    if (feeder_.joinable()) throw std::logic_error("BlockRing already started");
    pipeline.set_state(this);
    stopping_ = false;
    feeder_ = std::thread([this, &pipeline] { feed(pipeline); });
  }

  // Execute every submitted block, join the feeder and wait until the
  // pipeline has released every block handed to it (a GPUPipeline releases
  // each non-dummy block after its H2D copy, or when execute_pipeline fails),
  // so that no release arrives after the ring is gone.  Rethrows an exception
  // raised by execute_pipeline.
  void stop() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      stopping_ = true;
    }
    cv_ready_.notify_all();
    if (feeder_.joinable()) feeder_.join();
    {
      std::unique_lock<std::mutex> lock(mu_);
      cv_free_.wait(lock, [this] { return released_ >= executed_; });
    }
    if (error_) {
      auto e = error_;
      error_ = nullptr;
      std::rethrow_exception(e);
    }
  }

  // Called by the pipeline once the block's host memory is no longer read.
  void release_buffer(const size_t buffer_index) override {
    Block &b = block(buffer_index);
    std::memset(b.get_arrivals_ptr(), 0, b.get_arrivals_size());
    // notify under the lock: stop() may return (and the ring be destroyed) as
    // soon as it can see this release, which it cannot before the unlock
    std::lock_guard<std::mutex> lock(mu_);
    free_.push_back(buffer_index);
    ++released_;
    cv_free_.notify_all();
  }

  size_t blocks_executed() const {
    std::lock_guard<std::mutex> lock(mu_);
    return executed_;
  }
  size_t blocks_released() const {
    std::lock_guard<std::mutex> lock(mu_);
    return released_;
  }
  long missing_packets() const {
    std::lock_guard<std::mutex> lock(mu_);
    return missing_packets_;
  }

 private:
  void feed(GPUPipeline &pipeline) {
    for (;;) {
      size_t i;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_ready_.wait(lock, [this] { return stopping_ || !ready_.empty(); });
        if (ready_.empty()) return;  // stopping and drained
        i = ready_.front();
        ready_.pop_front();
      }
      try {
        pipeline.execute_pipeline(blocks_[i].get());
      } catch (...) {
        // the failed block returns through release_buffer (the pipeline
        // releases it once, as any handed block); the queued blocks go back to
        // the producer, so that a producer waiting in acquire() is not
        // stranded; stop() waits for the release and rethrows
        {
          std::lock_guard<std::mutex> lock(mu_);
          error_ = std::current_exception();
          failed_ = true;
          ++executed_;
          for (size_t j : ready_) free_.push_back(j);
          ready_.clear();
        }
        cv_free_.notify_all();
        return;
      }
      std::lock_guard<std::mutex> lock(mu_);
      ++executed_;
    }
  }

  std::vector<std::unique_ptr<Block>> blocks_;
  mutable std::mutex mu_;
  std::condition_variable cv_free_, cv_ready_;
  std::deque<size_t> free_, ready_;
  std::thread feeder_;
  bool stopping_ = false, failed_ = false;
  size_t executed_ = 0, released_ = 0;
  long missing_packets_ = 0;
  std::exception_ptr error_;
};
