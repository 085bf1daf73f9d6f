// Host worker threads for per-instance work of the batched steps.
//
// parallel_for(n, fn) runs fn(i) for i in [0, n) in contiguous chunks over up to
// 16 host threads (the calling thread takes the first chunk) and rethrows the
// first exception.  The threads are kept between calls: starting them for every
// call cost 0.26-0.43 ms, a tenth of a batched quasi-steady step.  One call uses
// the kept threads at a time; a call that finds them busy (another batch on
// another thread, or fn itself calling parallel_for), or a process forked from the
// one that started them, starts threads of its own for that call.
#pragma once

#include <unistd.h>

#include <algorithm>
#include <condition_variable>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace sb_pool {

class Workers {
public:
    // The threads wait on a condition variable between calls and end with the
    // process: the object is never destroyed (workers()), so neither the exit of this
    // process nor that of a process forked from it joins threads.
    explicit Workers(int n) : pid_(getpid()), n_threads_(n) {
        for (int w = 1; w <= n; ++w) std::thread([this, w] { loop(w); }).detach();
    }
    Workers(const Workers&) = delete;
    Workers& operator=(const Workers&) = delete;

    // fn(w) for w in [0, nt): w = 0 on the calling thread, the others on kept
    // threads.  Returns false (nothing run) when the kept threads are busy, too few,
    // or belong to another process.  fn must not throw (callers catch inside).
    bool run(int nt, const std::function<void(int)>& fn) {
        if (getpid() != pid_ || nt - 1 > n_threads_) return false;
        std::unique_lock<std::mutex> busy(run_m_, std::try_to_lock);
        if (!busy.owns_lock()) return false;
        {
            std::lock_guard<std::mutex> lk(m_);
            job_ = &fn;
            n_job_ = nt - 1;
            pending_ = nt - 1;
            ++gen_;
        }
        work_.notify_all();
        fn(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return pending_ == 0; });
        job_ = nullptr;
        return true;
    }

private:
    void loop(int w) {
        long seen = 0;
        for (;;) {
            const std::function<void(int)>* job = nullptr;
            {
                std::unique_lock<std::mutex> lk(m_);
                work_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (w > n_job_) continue;  // not needed for this call
                job = job_;
            }
            (*job)(w);
            std::lock_guard<std::mutex> lk(m_);
            if (--pending_ == 0) done_.notify_all();
        }
    }

    pid_t pid_;
    int n_threads_;
    std::mutex run_m_;  // held for a whole call
    std::mutex m_;
    std::condition_variable work_, done_;
    const std::function<void(int)>* job_ = nullptr;
    long gen_ = 0;
    int n_job_ = 0, pending_ = 0;
};

inline int max_threads() {
    unsigned hw = std::thread::hardware_concurrency();
    return (int)std::min<unsigned>(hw ? hw : 1, 16u);
}

inline Workers& workers() {
    static Workers* w = new Workers(max_threads() - 1);  // never destroyed
    return *w;
}

template <class F>
void parallel_for(int n, F&& fn) {
    int nt = std::min(max_threads(), std::max(1, n / 32));
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::exception_ptr> err(nt);
    const std::function<void(int)> chunk = [&](int w) {
        try {
            for (int i = (int)((long)n * w / nt); i < (int)((long)n * (w + 1) / nt); ++i) fn(i);
        } catch (...) {
            err[w] = std::current_exception();
        }
    };
    if (!workers().run(nt, chunk)) {
        std::vector<std::thread> pool;
        for (int w = 1; w < nt; ++w) pool.emplace_back(chunk, w);
        chunk(0);
        for (auto& th : pool) th.join();
    }
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

}  // namespace sb_pool
