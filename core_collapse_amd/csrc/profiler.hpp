// profiler.hpp -- per-kernel HIP-event timing on a handle's stream.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace cc {

class Profiler {
   public:
    bool on = false;

    // Record an event pair around one launch of kernel class `name` on stream `s`.
    void begin(hipStream_t s, const char* name) {
        if (!on) return;
        Rec r;
        r.name = name;
        r.a = get();
        r.b = get();
        hipEventRecord(r.a, s);
        open_.push_back(r);
    }
    void end(hipStream_t s) {
        if (!on || open_.empty()) return;
        Rec r = open_.back();
        open_.pop_back();
        hipEventRecord(r.b, s);
        done_.push_back(r);
    }
    // Fold finished pairs into the totals (the stream must have been synchronised).
    void collect() {
        for (auto& r : done_) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                auto& t = totals_[r.name];
                t.first += 1;
                t.second += ms;
            }
            pool_.push_back(r.a);
            pool_.push_back(r.b);
        }
        done_.clear();
    }
    void reset() {
        collect();
        totals_.clear();
    }
    bool get_total(const std::string& name, int* count, double* ms) {
        collect();
        auto it = totals_.find(name);
        if (it == totals_.end()) {
            *count = 0;
            *ms = 0.0;
            return false;
        }
        *count = it->second.first;
        *ms = it->second.second;
        return true;
    }
    ~Profiler() {
        for (auto& r : open_) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
        for (auto& r : done_) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
        for (auto e : pool_) hipEventDestroy(e);
    }

   private:
    struct Rec {
        const char* name;
        hipEvent_t a, b;
    };
    hipEvent_t get() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    std::vector<Rec> open_, done_;
    std::vector<hipEvent_t> pool_;
    std::map<std::string, std::pair<int, double>> totals_;
};

struct ProfScope {
    Profiler* p;
    hipStream_t s;
    ProfScope(Profiler* p_, hipStream_t s_, const char* name) : p(p_), s(s_) {
        if (p) p->begin(s, name);
    }
    ~ProfScope() {
        if (p) p->end(s);
    }
};

}  // namespace cc
