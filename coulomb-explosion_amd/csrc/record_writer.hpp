// The dataset generator's record writer: one thread writes each batch's records
// to the output file while the device computes the next batch and copies the
// one after.  Buffers are written in the order they are submitted;
// wait_written(k) returns once k of them are on their way to the file, so the
// caller may then refill the k-th buffer.  The caller owns the buffers and must
// not touch a submitted buffer before wait_written says it is done.
#pragma once

#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <fstream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>

namespace ce {

class RecordWriter {
 public:
  // Opens `path` for binary output; throws std::runtime_error if it cannot.
  explicit RecordWriter(const std::string& path) : out_(path, std::ios::binary) {
    if (!out_) throw std::runtime_error("cannot open output file: " + path);
    th_ = std::thread([this] { loop(); });
  }
  RecordWriter(const RecordWriter&) = delete;
  RecordWriter& operator=(const RecordWriter&) = delete;
  // A writer that was not finished stops after the queued buffers (errors are
  // dropped: call finish() to see them).
  ~RecordWriter() {
    if (th_.joinable()) {
      {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
      }
      cv_.notify_all();
      th_.join();
    }
  }

  void submit(const std::uint8_t* buf, std::size_t bytes) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      queue_.emplace_back(buf, bytes);
      ++submitted_;
    }
    cv_.notify_all();
  }
  void wait_written(std::size_t k) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [this, k] { return written_ >= k; });
  }
  std::size_t submitted() {
    std::lock_guard<std::mutex> lk(mu_);
    return submitted_;
  }
  // Writes what is queued, joins the thread and closes the file; throws
  // std::runtime_error if any write failed.
  void finish() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    th_.join();
    out_.close();
    if (!out_) throw std::runtime_error("writing the records failed");
  }

 private:
  void loop() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_.wait(lk, [this] { return !queue_.empty() || stop_; });
      if (queue_.empty()) return;  // stop_ with nothing left
      const auto [buf, bytes] = queue_.front();
      lk.unlock();
      out_.write(reinterpret_cast<const char*>(buf), static_cast<std::streamsize>(bytes));
      lk.lock();
      queue_.pop_front();
      ++written_;
      cv_.notify_all();
    }
  }

  std::ofstream out_;  // touched by the writer thread only, until it is joined
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::pair<const std::uint8_t*, std::size_t>> queue_;
  std::size_t submitted_ = 0, written_ = 0;
  bool stop_ = false;
  std::thread th_;
};

}  // namespace ce
