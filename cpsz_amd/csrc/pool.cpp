#include "pool.hpp"

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace cpsz {

class WorkerPool {
 public:
  WorkerPool(int device, int n) : device_(device) {
    for (int k = 0; k < n; k++) threads_.emplace_back([this] { loop(); });
  }
  int size() const { return (int)threads_.size(); }
  void run(std::vector<std::function<void(Ctx&)>>& jobs, int limit) {
    std::lock_guard<std::mutex> one_batch(run_m_);
    std::unique_lock<std::mutex> g(m_);
    limit_ = std::max(1, std::min(limit, size()));
    for (auto& j : jobs) q_.push_back(&j);
    pending_ = jobs.size();
    first_error_ = nullptr;
    cv_.notify_all();
    done_.wait(g, [&] { return pending_ == 0; });
    if (first_error_) std::rethrow_exception(first_error_);
  }

 private:
  void loop() {
    if (hipSetDevice(device_) != hipSuccess) return;
    std::unique_ptr<Ctx> ctx;
    for (;;) {
      std::function<void(Ctx&)>* job = nullptr;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return !q_.empty() && active_ < limit_; });
        job = q_.front();
        q_.pop_front();
        active_++;
      }
      std::exception_ptr err;
      try {
        if (!ctx) {
          ctx.reset(new Ctx(device_));
          ctx->batch_worker = true;
        }
        (*job)(*ctx);
      } catch (...) {
        err = std::current_exception();
      }
      {
        std::lock_guard<std::mutex> g(m_);
        active_--;
        if (err && !first_error_) first_error_ = err;
        if (--pending_ == 0) done_.notify_all();
      }
      cv_.notify_all();
    }
  }

  int device_;
  std::vector<std::thread> threads_;
  std::mutex run_m_, m_;
  std::condition_variable cv_, done_;
  std::deque<std::function<void(Ctx&)>*> q_;
  size_t pending_ = 0;
  int active_ = 0, limit_ = 1;
  std::exception_ptr first_error_;
};

WorkerPool& pool_for(int device) {
  static std::mutex m;
  static std::map<int, WorkerPool*>* pools = new std::map<int, WorkerPool*>();
  std::lock_guard<std::mutex> g(m);
  WorkerPool*& p = (*pools)[device];
  if (!p) p = new WorkerPool(device, kMaxWorkers);
  return *p;
}

void pool_run(WorkerPool& pool, std::vector<std::function<void(Ctx&)>>& jobs, int limit) {
  pool.run(jobs, limit);
}

struct HostHelper::State {
  std::mutex m;
  std::condition_variable cv;
  std::function<void()> job;
  bool pending = false, busy = false, quit = false;
  std::exception_ptr err;
  std::thread th;
};

HostHelper::HostHelper() : s_(new State) {
  State* s = s_;
  s->th = std::thread([s] {
    for (;;) {
      std::function<void()> job;
      {
        std::unique_lock<std::mutex> g(s->m);
        s->cv.wait(g, [&] { return s->pending || s->quit; });
        if (s->quit && !s->pending) return;
        job = std::move(s->job);
        s->pending = false;
      }
      std::exception_ptr err;
      try {
        job();
      } catch (...) {
        err = std::current_exception();
      }
      {
        std::lock_guard<std::mutex> g(s->m);
        s->err = err;
        s->busy = false;
      }
      s->cv.notify_all();
    }
  });
}

HostHelper::~HostHelper() {
  {
    std::lock_guard<std::mutex> g(s_->m);
    s_->quit = true;
  }
  s_->cv.notify_all();
  s_->th.join();
  delete s_;
}

void HostHelper::start(std::function<void()> job) {
  {
    std::lock_guard<std::mutex> g(s_->m);
    CPSZ_REQUIRE(!s_->busy, kErrInternal, "host helper already has a job");
    s_->job = std::move(job);
    s_->pending = s_->busy = true;
    s_->err = nullptr;
  }
  s_->cv.notify_all();
}

void HostHelper::wait() {
  std::unique_lock<std::mutex> g(s_->m);
  s_->cv.wait(g, [&] { return !s_->busy; });
  if (s_->err) {
    std::exception_ptr e = s_->err;
    s_->err = nullptr;
    std::rethrow_exception(e);
  }
}

HostHelper& host_helper() {
  thread_local HostHelper h;
  return h;
}

}  // namespace cpsz

