// Launch-shape autotuner.  Candidate launch shapes of a kernel are timed with
// HIP events on the caller's stream, on the caller's own data, and the fastest
// is kept per shape class for the rest of the process.  Classes are coarse
// (powers of two of the line and chunk counts, the FFT length, the profile and
// bin types), so one tuning serves every frame, simulator and batch job of that
// class.  Results do not depend on the choice beyond float summation order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <string>

namespace bcsim_hip {

// Tile-kernel launch shape; 0 in a field means "the heuristic's choice".
struct TileTune {
    int windows = 0, lines = 0, block = 0, slices = 0;
    int bins = 0;  // LDS bin accumulation: 1 fixed point, 2 float, 0 automatic
};

inline std::string describe(const TileTune& t) {
    char buf[112];
    std::snprintf(buf, sizeof buf, "lines=%d block=%d slices=%d windows=%d bins=%s", t.lines, t.block, t.slices,
                  t.windows, t.bins == 1 ? "fixed" : t.bins == 2 ? "float" : "auto");
    return buf;
}

class TuneCache {
  public:
    static TuneCache& get() {
        static TuneCache c;
        return c;
    }
    bool find_tile(const std::string& key, TileTune& out) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = tile_.find(key);
        if (it == tile_.end()) return false;
        out = it->second;
        return true;
    }
    void put_tile(const std::string& key, const TileTune& t, const std::string& record) {
        std::lock_guard<std::mutex> g(mu_);
        tile_[key] = t;
        records_[key] = record;
    }
    bool find_radix(const std::string& key, int& r) {
        std::lock_guard<std::mutex> g(mu_);
        auto it = radix_.find(key);
        if (it == radix_.end()) return false;
        r = it->second;
        return true;
    }
    void put_radix(const std::string& key, int r, const std::string& record) {
        std::lock_guard<std::mutex> g(mu_);
        radix_[key] = r;
        records_[key] = record;
    }
    // One line per tuned class: "key: choice (trial timings)".
    std::string report() {
        std::lock_guard<std::mutex> g(mu_);
        std::string s;
        for (auto& kv : records_) s += kv.first + ": " + kv.second + "\n";
        return s;
    }
    void clear() {
        std::lock_guard<std::mutex> g(mu_);
        tile_.clear();
        radix_.clear();
        records_.clear();
    }

  private:
    std::mutex mu_;
    std::map<std::string, TileTune> tile_;
    std::map<std::string, int> radix_;
    std::map<std::string, std::string> records_;
};

// Average duration in ms of `run` (which enqueues on `stream`): `warm` untimed
// launches, then `reps` launches between two events.
inline double time_launches(hipStream_t stream, const std::function<void()>& run, int reps, int warm = 2) {
    struct Ev {
        hipEvent_t e = nullptr;
        ~Ev() {
            if (e) (void)hipEventDestroy(e);
        }
    } a, b;
    for (int i = 0; i < warm; i++) run();
    if (hipEventCreate(&a.e) != hipSuccess || hipEventCreate(&b.e) != hipSuccess) return 0.0;
    (void)hipEventRecord(a.e, stream);
    for (int i = 0; i < reps; i++) run();
    (void)hipEventRecord(b.e, stream);
    float ms = 0.0f;
    if (hipEventSynchronize(b.e) == hipSuccess) (void)hipEventElapsedTime(&ms, a.e, b.e);
    return (double)ms / reps;
}

inline int ceil_log2(long long v) {
    int k = 0;
    while ((1LL << k) < v) k++;
    return k;
}

}  // namespace bcsim_hip
