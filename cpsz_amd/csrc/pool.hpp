// Worker threads shared by the batch drivers: each worker owns one context
// (HIP stream + scratch buffers) on the pool's device. Also the per-thread
// host helper the codecs use to overlap host work with their own.
#pragma once
#include <functional>
#include <vector>

#include "ctx.hpp"

namespace cpsz {

constexpr int kMaxWorkers = 16;

class WorkerPool;
// The device's pool of kMaxWorkers workers (created on first use, kept until
// the process ends: its threads and contexts must not be torn down after the
// HIP runtime at exit).
WorkerPool& pool_for(int device);
// Runs every job on at most `limit` workers at once and returns when all have
// finished; rethrows the first failure. Calls from different host threads
// are serialized.
void pool_run(WorkerPool& pool, std::vector<std::function<void(Ctx&)>>& jobs, int limit);

// A persistent host thread beside the calling thread, for host work that can
// run while the caller does its own (a second Huffman tree, the decode tables
// of a stream the device has not reached yet). Each calling thread has its
// own; the helper's thread-local storage persists from job to job.
class HostHelper {
 public:
  HostHelper();
  ~HostHelper();
  HostHelper(const HostHelper&) = delete;
  HostHelper& operator=(const HostHelper&) = delete;
  // Starts `job` (at most one job is in flight)
  void start(std::function<void()> job);
  // Waits for the job; rethrows its exception
  void wait();

 private:
  struct State;
  State* s_;
};
HostHelper& host_helper();

// A job on the calling thread's helper that is always waited for: the job
// uses objects of the caller's frame, so an exception leaving that frame
// first waits for it (and drops the job's own error). With `beside` false the
// job runs at once on the calling thread.
class HelperJob {
 public:
  HelperJob(std::function<void()> job, bool beside) : h_(beside ? &host_helper() : nullptr) {
    if (h_) h_->start(std::move(job));
    else job();
  }
  ~HelperJob() {
    if (h_ && !joined_) try {
        h_->wait();
      } catch (...) {
      }
  }
  void join() {
    joined_ = true;
    if (h_) h_->wait();
  }

 private:
  HostHelper* h_;
  bool joined_ = false;
};

}  // namespace cpsz
