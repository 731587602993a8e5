// A fixed set of host worker threads that run one job per call, all
// threads together, the caller being worker 0. Kept for the process by its
// users, so that a call does not start and join threads. A job catches its
// own exceptions: one that left a worker thread would end the process.
#pragma once
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace scarf {

class WorkerPool {
public:
    // A thread that cannot be started ends the construction with its
    // exception, after the threads already started are stopped and joined.
    explicit WorkerPool(int n) : n_(n)
    {
        try {
            for (int t = 1; t < n_; ++t) threads_.emplace_back([this, t] { loop(t); });
        } catch (...) {
            stop_all();
            throw;
        }
    }
    ~WorkerPool() { stop_all(); }
    int size() const { return n_; }
    void run(const std::function<void(int)> &job)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &job;
            pending_ = n_ - 1;
            ++generation_;
        }
        cv_.notify_all();
        job(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return pending_ == 0; });
        job_ = nullptr;
    }

private:
    void stop_all()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
        threads_.clear();
    }
    void loop(int t)
    {
        long seen = 0;
        for (;;) {
            const std::function<void(int)> *job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
                job = job_;
            }
            (*job)(t);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    int n_;
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *job_ = nullptr;
    long generation_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

}  // namespace scarf
