// Fixed-size thread pool with futures (the role of rvaser/thread_pool in
// racon: polisher.cpp:173-179, 369-378).
#pragma once

#include <condition_variable>
#include <cstdint>
#include <functional>
#include <future>
#include <mutex>
#include <queue>
#include <thread>
#include <unordered_map>
#include <vector>

namespace racon {

class ThreadPool {
 public:
  explicit ThreadPool(uint32_t num_threads) {
    if (num_threads == 0) {
      num_threads = 1;
    }
    for (uint32_t i = 0; i < num_threads; ++i) {
      threads_.emplace_back([this] { worker(); });
    }
    for (uint32_t i = 0; i < threads_.size(); ++i) {
      ids_[threads_[i].get_id()] = i;
    }
  }

  ~ThreadPool() {
    {
      std::lock_guard<std::mutex> lock(mutex_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) {
      t.join();
    }
  }

  uint32_t size() const { return static_cast<uint32_t>(threads_.size()); }

  // Index (0..size()-1) of the calling pool thread, or -1 from elsewhere.
  int32_t thread_index() const {
    auto it = ids_.find(std::this_thread::get_id());
    return it == ids_.end() ? -1 : static_cast<int32_t>(it->second);
  }

  template <typename F, typename... Args>
  auto submit(F&& f, Args&&... args)
      -> std::future<decltype(f(args...))> {
    using R = decltype(f(args...));
    auto task = std::make_shared<std::packaged_task<R()>>(
        std::bind(std::forward<F>(f), std::forward<Args>(args)...));
    std::future<R> result = task->get_future();
    {
      std::lock_guard<std::mutex> lock(mutex_);
      queue_.emplace([task] { (*task)(); });
    }
    cv_.notify_one();
    return result;
  }

 private:
  void worker() {
    while (true) {
      std::function<void()> job;
      {
        std::unique_lock<std::mutex> lock(mutex_);
        cv_.wait(lock, [this] { return stop_ || !queue_.empty(); });
        if (stop_ && queue_.empty()) {
          return;
        }
        job = std::move(queue_.front());
        queue_.pop();
      }
      job();
    }
  }

  std::vector<std::thread> threads_;
  std::unordered_map<std::thread::id, uint32_t> ids_;
  std::queue<std::function<void()>> queue_;
  std::mutex mutex_;
  std::condition_variable cv_;
  bool stop_ = false;
};

}  // namespace racon
