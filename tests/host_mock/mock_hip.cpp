// Test infrastructure (CPU only): a mock of the HIP runtime calls and of the
// kernel launchers that the library's host code uses, so that the host logic of
// the simulator, batch and server (sim_host.cpp, batch_host.cpp,
// serve_host.cpp) runs without a GPU and is checked for buffer lifetime and
// extent errors.  Device memory is host memory; streams run asynchronously on an
// executor thread in a random interleaving with random delays; every launch and
// copy checks, when it executes, that each buffer it touches lies inside a live
// allocation and that launch shapes fit the kernels' limits (LDS, list sizes,
// FFT extraction range).  Kernels compute nothing.  Elapsed times are random,
// so the launch-shape tuner picks random candidates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <set>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../openbcsim_amd/csrc/sim_device.h"  // launcher declarations and argument structs

using namespace bcsim_hip;

namespace {

struct Alloc {
    size_t size;
    bool pinned;
    uint64_t id;
};

std::mutex g_alloc_mu;
std::map<uintptr_t, Alloc> g_allocs;
std::map<uintptr_t, size_t> g_freed_pinned;  // host ranges that were pinned and are freed
std::atomic<uint64_t> g_alloc_id{1};
bool g_free_sync = getenv("MOCK_FREE_SYNC") ? atoi(getenv("MOCK_FREE_SYNC")) != 0 : true;
int g_max_delay_us = getenv("MOCK_DELAY_US") ? atoi(getenv("MOCK_DELAY_US")) : 300;

[[noreturn]] void die(const std::string& m) {
    fprintf(stderr, "\n*** MOCK CHECK FAILED: %s\n", m.c_str());
    fflush(stderr);
    abort();
}

// [p, p+n) inside one live allocation?  Unknown host pointers pass when allow_host.
void check(const void* p, size_t n, const char* what, const char* op, bool allow_host = false) {
    if (n == 0) return;
    const uintptr_t a = (uintptr_t)p;
    std::lock_guard<std::mutex> g(g_alloc_mu);
    auto it = g_allocs.upper_bound(a);
    if (it != g_allocs.begin()) {
        --it;
        if (a >= it->first && a < it->first + it->second.size) {
            if (a + n > it->first + it->second.size) {
                char b[512];
                snprintf(b, sizeof b, "%s: %s [%p, +%zu) overruns allocation %lu [%p, +%zu)", op, what, p, n,
                         (unsigned long)it->second.id, (void*)it->first, it->second.size);
                die(b);
            }
            return;
        }
    }
    if (allow_host) {
        auto f = g_freed_pinned.upper_bound(a);
        if (f != g_freed_pinned.begin()) {
            --f;
            if (a >= f->first && a < f->first + f->second) {
                char b[256];
                snprintf(b, sizeof b, "%s: %s %p lies in freed pinned memory", op, what, p);
                die(b);
            }
        }
        return;
    }
    char b[256];
    snprintf(b, sizeof b, "%s: %s %p (+%zu) is not in any live device allocation", op, what, p, n);
    die(b);
}

// ----------------------------------------------------------------- streams
struct Op {
    std::function<void()> fn;
    std::string name;
    ihipEvent_t* wait_ev = nullptr;  // hipStreamWaitEvent: runs once that record has completed
    uint64_t wait_gen = 0;
};

}  // namespace

struct ihipStream_t {
    int id;
    std::deque<Op> q;
    int running = 0;
    bool capturing = false;
    ihipGraph* graph = nullptr;
};
struct ihipEvent_t {
    // records enqueued / completed; an event is complete when both agree
    uint64_t recorded = 0;
    std::atomic<uint64_t> completed{0};
};
struct ihipGraph {
    std::vector<Op> ops;
};
struct hipGraphExec {
    std::vector<Op> ops;
};

namespace {

std::mutex g_q_mu;
std::condition_variable g_q_cv, g_done_cv;
std::vector<ihipStream_t*> g_streams;
ihipStream_t g_null{0};
std::atomic<int> g_stream_ids{1};
bool g_exec_started = false;
std::thread g_exec;
std::mt19937_64 g_rng(12345);

bool stream_ready(const ihipStream_t* s) {
    if (s->q.empty() || s->running) return false;
    const Op& op = s->q.front();
    return !op.wait_ev || op.wait_ev->completed.load() >= op.wait_gen;
}

void exec_loop() {
    std::unique_lock<std::mutex> lk(g_q_mu);
    for (;;) {
        g_q_cv.wait(lk, [] {
            for (auto* s : g_streams)
                if (stream_ready(s)) return true;
            return false;
        });
        std::vector<ihipStream_t*> ready;
        for (auto* s : g_streams)
            if (stream_ready(s)) ready.push_back(s);
        ihipStream_t* s = ready[g_rng() % ready.size()];
        Op op = std::move(s->q.front());
        s->q.pop_front();
        s->running = 1;
        const int delay = g_max_delay_us > 0 ? (int)(g_rng() % (uint64_t)g_max_delay_us) : 0;
        lk.unlock();
        if (delay) std::this_thread::sleep_for(std::chrono::microseconds(delay));
        op.fn();
        lk.lock();
        s->running = 0;
        g_done_cv.notify_all();
    }
}

void ensure_exec() {
    if (!g_exec_started) {
        g_exec_started = true;
        g_streams.push_back(&g_null);
        g_exec = std::thread(exec_loop);
        g_exec.detach();
    }
}

ihipStream_t* S(hipStream_t s) { return s ? s : &g_null; }

void enqueue(hipStream_t st, Op op) {
    std::lock_guard<std::mutex> g(g_q_mu);
    ensure_exec();
    ihipStream_t* s = S(st);
    if (s->capturing) {
        s->graph->ops.push_back(std::move(op));
        return;
    }
    s->q.push_back(std::move(op));
    g_q_cv.notify_all();
}

void drain(ihipStream_t* s) {
    std::unique_lock<std::mutex> lk(g_q_mu);
    g_done_cv.wait(lk, [&] { return s->q.empty() && s->running == 0; });
}

void drain_all() {
    std::unique_lock<std::mutex> lk(g_q_mu);
    g_done_cv.wait(lk, [&] {
        for (auto* s : g_streams)
            if (!s->q.empty() || s->running) return false;
        return true;
    });
}

}  // namespace

// ----------------------------------------------------------------- runtime API
extern "C" {

hipError_t hipMalloc(void** ptr, size_t size) {
    void* p = malloc(size ? size : 1);
    if (!p) return hipErrorOutOfMemory;
    std::lock_guard<std::mutex> g(g_alloc_mu);
    g_allocs[(uintptr_t)p] = Alloc{size, false, g_alloc_id++};
    *ptr = p;
    return hipSuccess;
}

hipError_t hipFree(void* ptr) {
    if (!ptr) return hipSuccess;
    if (g_free_sync) drain_all();
    {
        std::lock_guard<std::mutex> g(g_alloc_mu);
        auto it = g_allocs.find((uintptr_t)ptr);
        if (it == g_allocs.end() || it->second.pinned) die("hipFree of unknown pointer");
        g_allocs.erase(it);
    }
    free(ptr);
    return hipSuccess;
}

hipError_t hipHostMalloc(void** ptr, size_t size, unsigned int) {
    void* p = malloc(size ? size : 1);
    if (!p) return hipErrorOutOfMemory;
    std::lock_guard<std::mutex> g(g_alloc_mu);
    g_freed_pinned.erase((uintptr_t)p);
    g_allocs[(uintptr_t)p] = Alloc{size, true, g_alloc_id++};
    *ptr = p;
    return hipSuccess;
}

hipError_t hipHostFree(void* ptr) {
    if (!ptr) return hipSuccess;
    if (g_free_sync) drain_all();
    {
        std::lock_guard<std::mutex> g(g_alloc_mu);
        auto it = g_allocs.find((uintptr_t)ptr);
        if (it == g_allocs.end() || !it->second.pinned) die("hipHostFree of unknown pointer");
        g_freed_pinned[it->first] = it->second.size;
        g_allocs.erase(it);
    }
    // not returned to malloc: a later host buffer at this address would hide a use after free
    return hipSuccess;
}

hipError_t hipPointerGetAttributes(hipPointerAttribute_t* attr, const void* ptr) {
    std::lock_guard<std::mutex> g(g_alloc_mu);
    const uintptr_t p = (uintptr_t)ptr;
    auto it = g_allocs.upper_bound(p);
    if (it == g_allocs.begin()) return hipErrorInvalidValue;
    --it;
    if (p >= it->first + std::max<size_t>(it->second.size, 1)) return hipErrorInvalidValue;
    *attr = hipPointerAttribute_t{};
    attr->type = it->second.pinned ? hipMemoryTypeHost : hipMemoryTypeDevice;
    return hipSuccess;
}

hipError_t hipMemcpy(void* dst, const void* src, size_t n, hipMemcpyKind kind) {
    const bool dd = kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice;
    const bool sd = kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice;
    check(dst, n, "dst", "hipMemcpy", !dd);
    check(src, n, "src", "hipMemcpy", !sd);
    memcpy(dst, src, n);
    return hipSuccess;
}

hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind kind, hipStream_t stream) {
    const bool dd = kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice;
    const bool sd = kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice;
    enqueue(stream, Op{[=] {
                           check(dst, n, "dst", "hipMemcpyAsync", !dd);
                           check(src, n, "src", "hipMemcpyAsync", !sd);
                           memcpy(dst, src, n);
                       },
                       "memcpyAsync"});
    return hipSuccess;
}

hipError_t hipMemGetInfo(size_t* f, size_t* t) {
    *f = (size_t)200 << 30;
    *t = (size_t)288 << 30;
    return hipSuccess;
}

hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int* c) {
    *c = 1;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
    memset(p, 0, sizeof(*p));
    strcpy(p->name, "mock");
    strcpy(p->gcnArchName, "gfx950");
    p->multiProcessorCount = 256;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
    *v = a == hipDeviceAttributeMultiprocessorCount ? 256 : 0;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) {
    drain_all();
    return hipSuccess;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
    auto* st = new ihipStream_t();
    st->id = g_stream_ids++;
    std::lock_guard<std::mutex> g(g_q_mu);
    ensure_exec();
    g_streams.push_back(st);
    *s = st;
    return hipSuccess;
}

hipError_t hipStreamDestroy(hipStream_t s) {
    drain(s);
    std::lock_guard<std::mutex> g(g_q_mu);
    for (size_t i = 0; i < g_streams.size(); i++)
        if (g_streams[i] == s) {
            g_streams.erase(g_streams.begin() + i);
            break;
        }
    delete s;
    return hipSuccess;
}

hipError_t hipStreamSynchronize(hipStream_t s) {
    drain(S(s));
    return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new ihipEvent_t();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    uint64_t gen;
    {
        std::lock_guard<std::mutex> g(g_q_mu);
        gen = ++e->recorded;
    }
    enqueue(s, Op{[e, gen] {
                      uint64_t c = e->completed.load();
                      while (c < gen && !e->completed.compare_exchange_weak(c, gen)) {
                      }
                  },
                  "eventRecord"});
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    std::unique_lock<std::mutex> lk(g_q_mu);
    const uint64_t gen = e->recorded;
    g_done_cv.wait(lk, [&] { return e->completed.load() >= gen; });
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int flags) {
    if (flags != 0) die("hipStreamWaitEvent: flags must be 0");
    uint64_t gen;
    {
        std::lock_guard<std::mutex> g(g_q_mu);
        gen = e->recorded;
    }
    Op op{[] {}, "streamWaitEvent"};
    op.wait_ev = e;
    op.wait_gen = gen;
    enqueue(s, std::move(op));
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
    // random: the tuner picks random shapes.  The caller's thread has its own
    // engine (g_rng belongs to the executor thread, under g_q_mu).
    thread_local std::mt19937_64 rng(777);
    *ms = 0.01f + (float)(rng() % 1000) * 1e-3f;
    return hipSuccess;
}

hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "mock error"; }

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    std::lock_guard<std::mutex> g(g_q_mu);
    s->capturing = true;
    s->graph = new ihipGraph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    std::lock_guard<std::mutex> lk(g_q_mu);
    s->capturing = false;
    *g = s->graph;
    s->graph = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *e = new hipGraphExec();
    (*e)->ops = g->ops;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
    for (auto& op : e->ops) enqueue(s, op);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}

}  // extern "C"

// ----------------------------------------------------------------- kernel launchers
namespace bcsim_hip {

namespace {
constexpr size_t kLds = 160 * 1024;

void check_common(const ProjCommon& c, bool lut, const char* op) {
    check(c.lines, (size_t)c.num_lines * sizeof(LineGeom), "lines", op);
    if (lut) check(c.prof.quad, (size_t)c.prof.nr * c.prof.nl * c.prof.ne * 2 * sizeof(float4), "lut", op);
    if (lut && c.prof.slice_box) {
        // every level of the sparse table: nr entries each, levels while 2^k <= nr
        int levels = 1;
        while ((1 << levels) <= c.prof.nr) levels++;
        check(c.prof.slice_box, (size_t)levels * c.prof.nr * sizeof(int4), "slice boxes", op);
    }
}

void check_spline(const SplineSet& s, int num_lines, bool basis, const char* op) {
    const size_t cp = (size_t)s.num_cp * s.n * 4;
    if (s.iboxes) {
        if (s.n_iv < 1 || s.n_iv > s.num_cp) die(std::string(op) + ": knot interval count");
        check(s.iboxes, (size_t)((s.n + kChunk - 1) / kChunk) * s.n_iv * sizeof(ChunkBox), "interval boxes", op);
    }
    check(s.cx, cp, "cx", op);
    check(s.cy, cp, "cy", op);
    check(s.cz, cp, "cz", op);
    check(s.amp, (size_t)s.n * 4, "amp", op);
    if (basis) {
        check(s.basis, (size_t)num_lines * sizeof(SplineLineBasis), "basis", op);
        std::vector<SplineLineBasis> b(num_lines);
        memcpy(b.data(), s.basis, b.size() * sizeof(SplineLineBasis));
        for (auto& x : b)
            if (x.lo < 0 || x.count < 1 || x.lo + x.count > s.num_cp) die(std::string(op) + ": basis outside control points");
            else if (s.iboxes && x.lo >= s.n_iv) die(std::string(op) + ": basis interval outside the interval boxes");
    }
}

void check_tile(const ProjCommon& c, const FixedSet* fs, const SplineSet* ss, const ChunkBox* boxes, int nchunks,
                const float* bins, bool phase, bool lut, const TileLaunch& g, int block, const char* op) {
    const int E = phase ? 2 : 1;
    check_common(c, lut, op);
    if (ss) check_spline(*ss, c.num_lines, true, op);
    else check(fs->pts, (size_t)fs->n * 16, "pts", op);
    const int n = ss ? ss->n : fs->n;
    if (nchunks != (n + kChunk - 1) / kChunk) die(std::string(op) + ": nchunks does not match n");
    check(boxes, (size_t)nchunks * sizeof(ChunkBox), "boxes", op);
    check(bins, (size_t)c.num_lines * c.num_rf * E * 4, "bins", op);
    if (g.tile_lines < 1 || g.tile_lines > 32) die(std::string(op) + ": tile_lines");
    if (g.windows < 1 || (size_t)g.window_bins * g.windows < (size_t)c.num_rf ||
        (size_t)g.window_bins * (g.windows - 1) >= (size_t)c.num_rf)
        die(std::string(op) + ": windows do not tile the line");
    if ((size_t)g.slices * g.chunks_per_slice < (size_t)nchunks || g.slices < 1) die(std::string(op) + ": slices");
    if ((size_t)(g.slices - 1) * g.chunks_per_slice >= (size_t)nchunks) die(std::string(op) + ": empty slice");
    if (g.list_cap < 64 || g.list_cap > kMaxListCap || g.list_cap % 64) die(std::string(op) + ": list_cap");
    if (block < 64 || block > 1024 || block % 64) die(std::string(op) + ": block");
    if (g.list_packed && (g.tile_lines > 4 || nchunks >= (1 << 28))) die(std::string(op) + ": list_packed");
    const size_t lds =
        tile_lds_bytes(g.window_bins, tile_bin_bytes(phase, g.fixed), g.tile_lines, g.list_cap, g.list_packed);
    if (lds + 8 > kLds) die(std::string(op) + ": LDS over 160 KiB");
    if (g.fixed) {
        // fixed-point bins: a power-of-two scale and its exact inverse
        int e = 0;
        if (!(c.amp_scale > 0.0f) || std::frexp(c.amp_scale, &e) != 0.5f || c.amp_scale * c.inv_amp_scale != 1.0f)
            die(std::string(op) + ": fixed-point bins without a power-of-two amplitude scale");
    }
    if (g.chunks) {
        check(g.chunks, (size_t)nchunks * 4, "chunk order", op);
        std::vector<char> seen(nchunks, 0);
        for (int i = 0; i < nchunks; i++) {
            const int ci = g.chunks[i];
            if (ci < 0 || ci >= nchunks || seen[ci]) die(std::string(op) + ": chunk order is not a permutation");
            seen[ci] = 1;
        }
    }
    if (g.chunk_dist) {
        // the far and near bounds of the visiting order: ascending, after the order in its buffer
        if (!g.chunks || g.chunk_dist != reinterpret_cast<const float*>(g.chunks + nchunks))
            die(std::string(op) + ": chunk distances without their order");
        check(g.chunk_dist, (size_t)nchunks * 8, "chunk distances", op);
        for (int i = 1; i < 2 * nchunks; i++)
            if (i != nchunks && !(g.chunk_dist[i - 1] <= g.chunk_dist[i]))
                die(std::string(op) + ": chunk distance bounds not ascending");
        for (int i = 0; i < nchunks; i++)
            if (!(g.chunk_dist[nchunks + i] <= g.chunk_dist[i]) && std::isfinite(g.chunk_dist[i]))
                die(std::string(op) + ": near bound above far bound");
    }
    const int blocks = ((c.num_lines + g.tile_lines - 1) / g.tile_lines) * g.windows;
    if (g.work) check(g.work, (size_t)blocks * 4, "work", op);
    if (g.order) {
        check(g.order, (size_t)blocks * 4, "order", op);
        std::vector<char> seen(blocks, 0);
        for (int i = 0; i < blocks; i++) {
            const int w = g.order[i];
            if (w < 0 || w >= blocks || seen[w]) die(std::string(op) + ": order is not a permutation of the blocks");
            seen[w] = 1;
        }
    }
}

void check_conv(const ConvParams& p, const char* op) {
    const int E = p.complex_bins ? 2 : 1;
    check(p.bins, (size_t)p.num_lines * p.num_rf * E * 4, "bins", op);
    check(p.H, (size_t)p.fft_len * 8, "H", op);
    check(p.demod, (size_t)std::max(p.n_out, 1) * 8, "demod", op);
    const size_t tw = (size_t)(p.fft_len >> p.twiddle_bits) + ((size_t)1 << p.twiddle_bits);
    check(p.twiddles, tw * 8, "twiddles", op);
    check(p.iq, (size_t)p.num_lines * p.n_out * 8, "iq", op);
    if (p.noise_counter_ptr) check(p.noise_counter_ptr, 4, "noise counter", op);
    if (p.fft_len > kMaxFftLen || (1 << p.log2_len) != p.fft_len) die(std::string(op) + ": fft_len");
    if (p.twiddle_bits != p.log2_len / 2) die(std::string(op) + ": twiddle bits");
    if (p.n_out > 0 && (long long)p.center + (long long)(p.n_out - 1) * p.decimation >= p.fft_len)
        die(std::string(op) + ": extraction beyond the transform");
    if (p.radix != 4 && p.radix != 8) die(std::string(op) + ": radix");
    if (p.fold_log2 < 0 || p.log2_len - p.fold_log2 < 4 || (p.out_stride << p.fold_log2) != p.decimation)
        die(std::string(op) + ": fold");
    if (p.n_out > 0 && (long long)(p.n_out - 1) * p.out_stride >= (p.fft_len >> p.fold_log2))
        die(std::string(op) + ": extraction beyond the folded transform");
    if (p.num_rf > p.fft_len) die(std::string(op) + ": num_rf > fft_len");
}
}  // namespace

size_t tile_lds_bytes(int window_bins, size_t bin_bytes, int tile_lines, int list_cap, int list_packed) {
    return (size_t)tile_lines * ((size_t)window_bins * bin_bytes + kTileLineLds) +
           (size_t)list_cap * (list_packed ? 4 : 8);
}
int conv_block_threads(int fft_len) { return fft_len >= 4096 ? 1024 : 256; }

hipError_t launch_fill_zero(float* p, size_t n, hipStream_t s) {
    enqueue(s, Op{[=] { check(p, n * 4, "bins", "fill_zero"); }, "fill_zero"});
    return hipSuccess;
}

hipError_t launch_project_fixed(const ProjCommon& c, const FixedSet& fs, void* bins, bool cplx, bool, bool,
                                bool lut, const ProjLaunch& cfg, hipStream_t s) {
    enqueue(s, Op{[=] {
                      check_common(c, lut, "project_fixed");
                      check(fs.pts, (size_t)fs.n * 16, "pts", "project_fixed");
                      check(bins, (size_t)c.num_lines * c.num_rf * (cplx ? 8 : 4), "bins", "project_fixed");
                      if ((size_t)cfg.tile_lines * c.num_rf * (cplx ? 8 : 4) > kLds) die("project_fixed LDS");
                  },
                  "project_fixed"});
    return hipSuccess;
}

hipError_t launch_project_spline(const ProjCommon& c, const SplineSet& ss, void* bins, bool cplx, bool, bool,
                                 bool lut, const ProjLaunch& cfg, hipStream_t s) {
    enqueue(s, Op{[=] {
                      check_common(c, lut, "project_spline");
                      check_spline(ss, c.num_lines, true, "project_spline");
                      check(bins, (size_t)c.num_lines * c.num_rf * (cplx ? 8 : 4), "bins", "project_spline");
                      if ((size_t)cfg.tile_lines * c.num_rf * (cplx ? 8 : 4) > kLds) die("project_spline LDS");
                  },
                  "project_spline"});
    return hipSuccess;
}

hipError_t launch_render_spline(const SplineSet& ss, const SplineLineBasis& b, float4* out, hipStream_t s) {
    enqueue(s, Op{[=] {
                      check_spline(ss, 0, false, "render");
                      if (b.lo < 0 || b.lo + b.count > ss.num_cp) die("render: basis outside control points");
                      check(out, (size_t)ss.n * 16, "rendered", "render");
                  },
                  "render"});
    return hipSuccess;
}

hipError_t launch_convolve_demod(const ConvParams& p, hipStream_t s) {
    enqueue(s, Op{[=] { check_conv(p, "convolve"); }, "convolve"});
    return hipSuccess;
}

hipError_t launch_project_tile(const ProjCommon& c, const CullParams&, const FixedSet* fixed, const SplineSet* spline,
                               const ChunkBox* boxes, int nchunks, void* bins, bool, bool phase, bool lut, bool,
                               const TileLaunch& cfg, hipStream_t s) {
    if (nchunks == 0) return hipSuccess;
    const bool is_spline = spline != nullptr;
    const FixedSet fs = fixed ? *fixed : FixedSet{nullptr, 0};
    const SplineSet ss = spline ? *spline : SplineSet{};
    enqueue(s, Op{[=] {
                      check_tile(c, is_spline ? nullptr : &fs, is_spline ? &ss : nullptr, boxes, nchunks,
                                 (const float*)bins, phase, lut, cfg, cfg.block, "project_tile");
                      if (cfg.stats) check(cfg.stats, 3 * sizeof(unsigned long long), "stats", "project_tile");
                  },
                  "project_tile"});
    return hipSuccess;
}

hipError_t launch_order_tiles(int* work, int* order, int n, hipStream_t s) {
    if (n <= 0 || n > kMaxOrderedTiles) return hipErrorInvalidValue;
    enqueue(s, Op{[=] {
                      check(work, (size_t)n * 4, "work", "order_tiles");
                      check(order, (size_t)n * 4, "order", "order_tiles");
                      std::vector<int> idx(n);
                      for (int i = 0; i < n; i++) idx[i] = i;
                      std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return work[a] > work[b]; });
                      for (int i = 0; i < n; i++) order[i] = idx[i], work[i] = 0;
                  },
                  "order_tiles"});
    return hipSuccess;
}

hipError_t launch_order_segments(const OrderSeg* segs, int n_segs, hipStream_t s) {
    if (n_segs <= 0) return hipSuccess;
    enqueue(s, Op{[=] {
                      check(segs, (size_t)n_segs * sizeof(OrderSeg), "segments", "order_segments");
                      for (int k = 0; k < n_segs; k++) {
                          const OrderSeg& S = segs[k];
                          if (S.n <= 0 || S.n > kMaxOrderedTiles) die("order_segments: segment size");
                          if (S.slices < 1) die("order_segments: slices");
                          check(S.work, (size_t)S.n * 4, "work", "order_segments");
                          check(S.items, (size_t)S.n * S.slices * sizeof(int2), "items", "order_segments");
                          const int job = S.items[0].x;
                          for (size_t i = 0; i < (size_t)S.n * S.slices; i++)
                              if (S.items[i].x != job) die("order_segments: items of another job in the segment");
                          std::vector<int> idx(S.n);
                          for (int i = 0; i < S.n; i++) idx[i] = i;
                          std::stable_sort(idx.begin(), idx.end(),
                                           [&](int a, int b) { return S.work[a] > S.work[b]; });
                          for (int r = 0; r < S.n; r++)
                              for (int sl = 0; sl < S.slices; sl++) S.items[(size_t)sl * S.n + r].y = sl * S.n + idx[r];
                          for (int i = 0; i < S.n; i++) S.work[i] = 0;
                      }
                  },
                  "order_segments"});
    return hipSuccess;
}

hipError_t launch_project_batch(int variant, const ProjJob* jobs, const int2* items, int n_items, int* counter,
                                size_t lds, int, hipStream_t s) {
    if (n_items == 0) return hipSuccess;
    enqueue(s, Op{[=] {
                      const char* op = "project_batch";
                      check(counter, 4, "counter", op);
                      check(items, (size_t)n_items * sizeof(int2), "items", op);
                      int nj = 0;
                      for (int i = 0; i < n_items; i++) nj = std::max(nj, items[i].x + 1);
                      check(jobs, (size_t)nj * sizeof(ProjJob), "jobs", op);
                      for (int i = 0; i < n_items; i++)
                          if (batch_variant(jobs[items[i].x].flags) != variant) die("project_batch: item of another variant");
                      std::set<int> present;  // the jobs of this launch's variant
                      for (int i = 0; i < n_items; i++) present.insert(items[i].x);
                      for (int j : present) {
                          const ProjJob& J = jobs[j];
                          const bool spl = J.flags & kJobSpline;
                          check_tile(J.c, spl ? nullptr : &J.fs, spl ? &J.ss : nullptr, J.boxes, J.nchunks, J.bins,
                                     J.flags & kJobPhase, J.flags & kJobLut, J.g, 1024, op);
                          const int tiles = (J.c.num_lines + J.g.tile_lines - 1) / J.g.tile_lines;
                          if (J.blocks_per_slice != tiles * J.g.windows) die("project_batch: blocks_per_slice");
                          if (tile_lds_bytes(J.g.window_bins, tile_bin_bytes(J.flags & kJobPhase, J.g.fixed),
                                             J.g.tile_lines, J.g.list_cap, J.g.list_packed) > lds)
                              die("project_batch: job needs more LDS than launched");
                          if (((J.flags & kJobFixed) != 0) != (J.g.fixed != 0))
                              die("project_batch: job flags and launch disagree on fixed-point bins");
                      }
                      std::set<std::pair<int, int>> seen;
                      for (int i = 0; i < n_items; i++) {
                          const ProjJob& J = jobs[items[i].x];
                          if (items[i].y < 0 || items[i].y >= J.blocks_per_slice * J.g.slices) die("project_batch: item");
                          if (!seen.insert({items[i].x, items[i].y}).second) die("project_batch: item listed twice");
                      }
                      size_t all = 0;
                      for (int j : present) all += (size_t)jobs[j].blocks_per_slice * jobs[j].g.slices;
                      if (all != (size_t)n_items) die("project_batch: items do not cover every block");
                      if (lds + 8 > kLds) die("project_batch: LDS");
                  },
                  "project_batch"});
    return hipSuccess;
}

hipError_t launch_convolve_batch(int variant, const ConvParams* jobs, const int2* items, int n_items, size_t lds,
                                 hipStream_t s) {
    if (n_items == 0) return hipSuccess;
    enqueue(s, Op{[=] {
                      const char* op = "convolve_batch";
                      check(items, (size_t)n_items * sizeof(int2), "items", op);
                      if (lds > kLds) die("convolve_batch: LDS over 160 KiB");
                      int nj = 0;
                      for (int i = 0; i < n_items; i++) nj = std::max(nj, items[i].x + 1);
                      check(jobs, (size_t)nj * sizeof(ConvParams), "jobs", op);
                      std::set<int> present;
                      for (int i = 0; i < n_items; i++) present.insert(items[i].x);
                      for (int j : present) {
                          const ConvParams& p = jobs[j];
                          check_conv(p, op);
                          if (conv_variant(p, conv_block_threads(p.fft_len)) != variant)
                              die("convolve_batch: item of another variant");
                          if (conv_lds_bytes(p) > lds) die("convolve_batch: job needs more LDS than launched");
                      }
                      for (int i = 0; i < n_items; i++)
                          if (items[i].y < 0 || items[i].y >= jobs[items[i].x].num_lines) die("convolve_batch: item");
                  },
                  "convolve_batch"});
    return hipSuccess;
}

// The library's fixed scratch array of the allocation-free frame: one mock device
// allocation, looked up like hipGetSymbolAddress would.
hipError_t noalloc_scratch(int device, void** ptr) {
    static std::mutex mu;
    static void* addr = nullptr;
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!addr && hipMalloc(&addr, kNoAllocScratchBytes) != hipSuccess) return hipErrorOutOfMemory;
    *ptr = addr;
    return hipSuccess;
}

hipError_t launch_fill_segments(const FillSeg* segs, int n, int* counters, int nc, uint32_t* noise, int n_noise,
                                hipStream_t s) {
    enqueue(s, Op{[=] {
                      check(counters, (size_t)nc * 4, "counters", "fill_segments");
                      check(noise, (size_t)n_noise * 4, "noise counters", "fill_segments");
                      if (n) check(segs, (size_t)n * sizeof(FillSeg), "segs", "fill_segments");
                      for (int i = 0; i < n; i++) check(segs[i].p, segs[i].n * 4, "segment", "fill_segments");
                  },
                  "fill_segments"});
    return hipSuccess;
}

}  // namespace bcsim_hip
