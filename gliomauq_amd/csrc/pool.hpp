// pool.hpp -- host worker threads for the short data-parallel loops of an
// adaptation (table builds, planning).  Such a loop takes tens of microseconds
// and the loops of one adaptation follow each other closely, so the workers
// spin between jobs (a wake-up from sleep would cost as much as the loop) and
// go to sleep only after a quiet period.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace gq {

class WorkerPool {
public:
    explicit WorkerPool(int nworkers)
    {
        // a worker starts from generation 0, so it also sees a job or the
        // stop published before it first ran
        for (int i = 0; i < nworkers; i++) th_.emplace_back([this] { work(0u); });
    }
    ~WorkerPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_.store(true, std::memory_order_release);
            gen_.fetch_add(1, std::memory_order_acq_rel);
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    WorkerPool(const WorkerPool &) = delete;
    WorkerPool &operator=(const WorkerPool &) = delete;

    int workers() const { return (int)th_.size(); }

    // f(begin, end) for consecutive chunks of [0, n) of `grain` items; the
    // calling thread takes chunks too (and runs f(0, n) itself when the pool
    // has no workers or n fits one chunk).  Returns when every chunk is done
    // (no chunk runs after that); the first exception thrown by f is rethrown.
    void run(int n, int grain, const std::function<void(int, int)> &f)
    {
        if (n <= 0) return;
        if (grain < 1) grain = 1;
        const int nchunks = (n + grain - 1) / grain;
        if (th_.empty() || nchunks == 1) {
            f(0, n);
            return;
        }
        // each job is its own object: a worker that arrives late finds its
        // chunks taken and leaves without touching a newer job's state
        auto job = std::make_shared<Job>();
        job->f = &f;
        job->n = n;
        job->grain = grain;
        job->nchunks = nchunks;
        std::atomic_store(&cur_, job);
        {
            std::lock_guard<std::mutex> lk(mu_);
            gen_.fetch_add(1, std::memory_order_acq_rel);
        }
        if (sleepers_.load(std::memory_order_acquire) > 0) cv_.notify_all();
        take(*job);
        while (job->done.load(std::memory_order_acquire) < nchunks) std::this_thread::yield();
        std::atomic_store(&cur_, std::shared_ptr<Job>());
        if (job->err) std::rethrow_exception(job->err);
    }

private:
    struct Job {
        const std::function<void(int, int)> *f = nullptr;
        int n = 0, grain = 1, nchunks = 0;
        std::atomic<int> next{0}, done{0};
        std::mutex emu;
        std::exception_ptr err;
    };

    static void take(Job &j)
    {
        for (int c; (c = j.next.fetch_add(1, std::memory_order_acq_rel)) < j.nchunks;) {
            const int b = c * j.grain, e = b + j.grain < j.n ? b + j.grain : j.n;
            try {
                (*j.f)(b, e);
            } catch (...) {
                std::lock_guard<std::mutex> lk(j.emu);
                if (!j.err) j.err = std::current_exception();
            }
            j.done.fetch_add(1, std::memory_order_acq_rel);
        }
    }

    void work(unsigned seen)
    {
        for (;;) {
            // spin for a while, then sleep until the next job (or the stop)
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load(std::memory_order_acquire) == seen && !stop_.load(std::memory_order_acquire)) {
                if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
                    std::unique_lock<std::mutex> lk(mu_);
                    sleepers_.fetch_add(1, std::memory_order_acq_rel);
                    cv_.wait(lk, [&] {
                        return gen_.load(std::memory_order_acquire) != seen || stop_.load(std::memory_order_acquire);
                    });
                    sleepers_.fetch_sub(1, std::memory_order_acq_rel);
                    break;
                }
                std::this_thread::yield();
            }
            seen = gen_.load(std::memory_order_acquire);
            if (stop_.load(std::memory_order_acquire)) return;
            const std::shared_ptr<Job> job = std::atomic_load(&cur_);
            if (job) take(*job);
        }
    }

    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::atomic<unsigned> gen_{0};
    std::atomic<int> sleepers_{0};
    std::atomic<bool> stop_{false};
    std::shared_ptr<Job> cur_;
};

}  // namespace gq
