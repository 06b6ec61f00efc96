// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A small fixed pool of host threads for the oracle's data-parallel loops (per
// body, per pair, per manifold, per colour batch).  Every loop hands out fixed
// index ranges and every reduction combines per-range partial results in range
// order, so a step gives bit-identical results for any thread count; the pool
// only changes how fast the oracle runs (the all-core CPU baseline).
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

class Pool {
public:
    explicit Pool(unsigned threads = 1) { resize(threads); }
    ~Pool() { stop(); }
    Pool(const Pool&) = delete;
    Pool& operator=(const Pool&) = delete;

    unsigned size() const { return (unsigned)workers_.size() + 1u; }

    void resize(unsigned threads) {
        stop();
        if (threads < 1) threads = 1;
        quit_ = false;
        for (unsigned t = 1; t < threads; ++t) workers_.emplace_back([this]() { loop(); });
    }

    // Calls fn(c, begin, end) for the chunks c = 0 .. chunks-1 of [0, n), ranges of
    // (nearly) equal size; they depend on n and chunks only, never on the thread count.
    void run(uint64_t n, uint64_t chunks, const std::function<void(uint64_t, uint64_t, uint64_t)>& fn) {
        if (n == 0) return;
        if (chunks < 1) chunks = 1;
        if (chunks > n) chunks = n;
        if (workers_.empty() || chunks == 1) {
            for (uint64_t c = 0; c < chunks; ++c) fn(c, n * c / chunks, n * (c + 1) / chunks);
            return;
        }
        Job job;
        job.fn = &fn;
        job.n = n;
        job.chunks = chunks;
        {
            std::lock_guard<std::mutex> lk(m_);
            cur_ = &job;
            ++gen_;
        }
        cv_.notify_all();
        work(job);
        while (job.done.load(std::memory_order_acquire) < chunks) std::this_thread::yield();
        {
            // no worker picks the job up after this; wait for those that did to let go of it
            std::lock_guard<std::mutex> lk(m_);
            cur_ = nullptr;
        }
        while (active_.load(std::memory_order_acquire) != 0) std::this_thread::yield();
    }

    // The chunk count parallel_for uses for n items of grain `grain`.
    static uint64_t chunks_for(uint64_t n, uint64_t grain) {
        uint64_t c = grain ? (n + grain - 1) / grain : 1;
        if (c > 4096) c = 4096;
        if (c > n) c = n;
        return c < 1 ? 1 : c;
    }
    // Loop over [0, n) in chunks_for(n, grain) ranges.
    void parallel_for(uint64_t n, uint64_t grain, const std::function<void(uint64_t, uint64_t, uint64_t)>& fn) {
        run(n, chunks_for(n, grain), fn);
    }

private:
    struct Job {
        const std::function<void(uint64_t, uint64_t, uint64_t)>* fn = nullptr;
        uint64_t n = 0, chunks = 0;
        std::atomic<uint64_t> next{0}, done{0};
    };
    static void work(Job& j) {
        for (;;) {
            const uint64_t c = j.next.fetch_add(1);
            if (c >= j.chunks) break;
            (*j.fn)(c, j.n * c / j.chunks, j.n * (c + 1) / j.chunks);
            j.done.fetch_add(1, std::memory_order_acq_rel);
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            Job* j = nullptr;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&]() { return quit_ || (gen_ != seen && cur_ != nullptr); });
                if (quit_) return;
                seen = gen_;
                j = cur_;
                active_.fetch_add(1);
            }
            work(*j);
            active_.fetch_sub(1, std::memory_order_acq_rel);
        }
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(m_);
            quit_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
        workers_.clear();
    }

    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_;
    Job* cur_ = nullptr;
    uint64_t gen_ = 0;
    std::atomic<uint32_t> active_{0};
    bool quit_ = false;
};

}  // namespace orc
