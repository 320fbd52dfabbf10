// Per-kernel device timing with HIP events on one stream: each timed launch
// is bracketed by an event pair from a reusable pool; pairs are resolved
// after a stream synchronisation.  Also accumulates the launch's
// algorithmic byte count so that achieved bandwidth is bytes / time.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"

namespace owl {

class Profiler {
public:
    static constexpr int kKinds = 16;
    ~Profiler() {
        for (auto e : pool_) (void)hipEventDestroy(e);
    }
    void enable(bool on) {
        on_ = on;
        for (int k = 0; k < kKinds; ++k) ms_[k] = bytes_[k] = 0, launches_[k] = 0;
        used_.clear();
        next_ = 0;
    }
    bool on() const { return on_; }
    // At most kMaxPending pairs stay in flight: beyond that the stream is
    // synchronised and the recorded pairs are folded into the totals, so a
    // long run holds a bounded number of events.
    static constexpr size_t kMaxPending = 4096;
    int begin(hipStream_t s, int kind, double bytes) {
        if (!on_) return -1;
        if (used_.size() >= kMaxPending) {
            OWL_HIP(hipStreamSynchronize(s));
            collect();
        }
        if (next_ + 2 > pool_.size())
            for (int k = 0; k < 256; ++k) {
                hipEvent_t e;
                OWL_HIP(hipEventCreate(&e));
                pool_.push_back(e);
            }
        const int first = int(next_);
        next_ += 2;
        OWL_HIP(hipEventRecord(pool_[first], s));
        used_.push_back(Use{kind, first, bytes});
        return first;
    }
    void end(hipStream_t s, int first) {
        if (first >= 0) OWL_HIP(hipEventRecord(pool_[first + 1], s));
    }
    // Call after the stream has been synchronised.
    void collect() {
        for (auto& u : used_) {
            float ms = 0.f;
            OWL_HIP(hipEventElapsedTime(&ms, pool_[u.first], pool_[u.first + 1]));
            ms_[u.kind] += ms;
            bytes_[u.kind] += u.bytes;
            launches_[u.kind] += 1;
        }
        used_.clear();
        next_ = 0;
    }
    size_t pending() const { return used_.size(); }
    double ms(int k) const { return ms_[k]; }
    double bytes(int k) const { return bytes_[k]; }
    long long launches(int k) const { return launches_[k]; }

private:
    struct Use {
        int kind, first;
        double bytes;
    };
    bool on_ = false;
    std::vector<hipEvent_t> pool_;
    std::vector<Use> used_;
    size_t next_ = 0;
    double ms_[kKinds] = {0}, bytes_[kKinds] = {0};
    long long launches_[kKinds] = {0};
};

}  // namespace owl
