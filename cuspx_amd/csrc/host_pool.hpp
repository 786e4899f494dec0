// host_pool.hpp -- a few persistent host threads for the host-buffer calls'
// copies between the caller's buffers and pinned staging.  Standard C++
// only (no HIP), so a native test can build it under the sanitizers.
#pragma once
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

namespace spx {

// Host threads that move data between the caller's buffers and the pinned
// staging: one thread copies ~8-10 GB/s, so a 4096-signature batch (70 MB)
// would otherwise spend longer in memcpy than on the GPU.  run(n, fn) calls
// fn(0) .. fn(n-1), each once, on the workers and the calling thread; indices
// are claimed in increasing order, and run returns when every call has.  The
// pool lives until the process ends (never destroyed, so no thread is joined
// during static destruction).
class HostPool {
  public:
    static HostPool& get() {
        static HostPool* p = new HostPool();
        return *p;
    }

    void run(uint32_t n, const std::function<void(uint32_t)>& fn) {
        if (n == 0) return;
        if (n == 1 || workers_ == 0) {
            for (uint32_t k = 0; k < n; k++) fn(k);
            return;
        }
        std::lock_guard<std::mutex> one(call_mu_);  // one batch at a time
        Batch b{&fn, n};
        {
            std::lock_guard<std::mutex> lk(mu_);
            cur_ = &b;
            epoch_++;
        }
        cv_.notify_all();
        drain(b);
        std::unique_lock<std::mutex> lk(mu_);
        cur_ = nullptr;  // late wakers find nothing; wait for those inside b
        done_.wait(lk, [&] { return busy_ == 0; });
    }

  private:
    struct Batch {
        const std::function<void(uint32_t)>* fn;
        uint32_t n;
        std::atomic<uint32_t> next{0};
    };

    HostPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        workers_ = hw > 1 ? (hw < 8 ? hw : 8) - 1 : 0;
        for (unsigned i = 0; i < workers_; i++) std::thread([this] { loop(); }).detach();
    }

    static void drain(Batch& b) {
        for (uint32_t k = b.next.fetch_add(1); k < b.n; k = b.next.fetch_add(1)) (*b.fn)(k);
    }

    void loop() {
        uint64_t seen = 0;
        for (;;) {
            Batch* b;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return cur_ && epoch_ != seen; });
                seen = epoch_;
                b = cur_;
                busy_++;
            }
            drain(*b);
            std::lock_guard<std::mutex> lk(mu_);
            if (--busy_ == 0) done_.notify_all();
        }
    }

    unsigned workers_ = 0;
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    Batch* cur_ = nullptr;
    uint64_t epoch_ = 0;
    unsigned busy_ = 0;
};

}  // namespace spx
