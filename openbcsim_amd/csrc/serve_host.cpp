// Serving loop: frame requests against resident phantoms, admitted by a device
// memory budget and batched continuously.
//
// The reference serves one frame per simulate_lines call from the caller's
// thread (GpuAlgorithm.cpp:189-360; the GUI's packet loop, MainWindow.cpp:666-695,
// reconfigures the scan and calls again).  Here a server holds phantoms
// (snapshots of configured simulators whose datasets and LUT stay in HBM) and
// scan geometries, and takes frame requests (phantom, geometry, time t) from
// any thread:
//   - a scheduler thread admits pending requests in arrival order while their
//     frame buffers fit the memory budget (buffers of finished requests are kept
//     for the next request of the same phantom and geometry, or evicted least
//     recently used first), and launches the admitted requests as one batched
//     frame (one persistent projection launch, one persistent convolution
//     launch) on one of `contexts` frame streams, so the next batches are
//     formed and launched while earlier ones compute or copy;
//   - a request of more than `piece_lines` lines is served in pieces
//     (consecutive line ranges), at most one piece of a request per frame, and
//     frames holding pieces run their kernels one after another: the IQ of one
//     piece crosses PCIe while the next piece computes, instead of the whole
//     frame's copy starting after its last kernel;
//   - at the end of its frame each request's IQ is copied by DMA, smallest
//     first, straight into the caller's buffer when that is pinned memory of the
//     server's (bcsim_server_alloc_host), otherwise into a pinned staging
//     buffer; a completion thread completes the requests one by one as their
//     copies land, and releases each frame's buffers, which admits waiting
//     requests;
//   - for a staged request the caller's wait copies the IQ into the caller's
//     buffer, in the caller's thread (several threads for large frames), so one
//     large result does not hold up the completion of the frames behind it.
// Each request has its own bins and tables, so its result does not depend on
// what it is batched with, up to float summation order.
#include "batch_state.h"
#include "host_copy.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <set>
#include <thread>

namespace bcsim_hip {
namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Geometry {
    std::vector<float> o, d, lat, t;
    float length = 0.f;
    int lines() const { return (int)t.size(); }
    // consecutive line ranges of at most `piece_size` lines, each an internal
    // geometry (ids of bcsim_server::piece_geoms) served as a job of its own
    int piece_size = 0;
    std::vector<int> pieces;
};

// Frame buffers of one in-flight or idle request: a simulator sharing its
// phantom's datasets, with the geometry's scan.
struct Slot {
    std::unique_ptr<bcsim_sim> sim;
    int phantom = -1, geometry = -1;
    size_t bytes = 0;  // accounted device bytes
    bool busy = false;
    uint64_t last_used = 0;
};

// Pinned host buffers by power-of-two size class, reused across requests
// (pinned allocation is slow and may stall the device).
class PinnedPool {
  public:
    ~PinnedPool() {
        for (auto& kv : free_)
            for (void* p : kv.second) (void)hipHostFree(p);
    }
    void* take(size_t bytes, size_t& cls) {
        cls = 4096;
        while (cls < bytes) cls <<= 1;
        {
            std::lock_guard<std::mutex> g(mu_);
            auto& v = free_[cls];
            if (!v.empty()) {
                void* p = v.back();
                v.pop_back();
                pooled_ -= cls;
                return p;
            }
        }
        void* p = nullptr;
        hip_check(hipHostMalloc(&p, cls, 0), "hipHostMalloc");
        return p;
    }
    void give(void* p, size_t cls) {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        if (pooled_ + cls > kMaxPooled) {
            (void)hipHostFree(p);
            return;
        }
        free_[cls].push_back(p);
        pooled_ += cls;
    }

  private:
    static constexpr size_t kMaxPooled = (size_t)8 << 30;
    std::mutex mu_;
    std::map<size_t, std::vector<void*>> free_;
    size_t pooled_ = 0;
};

struct Request {
    uint64_t id = 0;
    int phantom = 0, geometry = 0;
    float t = 0.f;
    float* out = nullptr;  // caller's complex64 [lines][samples]
    size_t capacity = 0;
    int lines = 0, samples = 0;
    double t_submit = 0, t_launch = 0, t_done = 0;  // t_done: IQ staged in pinned memory
    int status = BCSIM_OK;
    std::string error;
    bool done = false;
    Slot* slot = nullptr;
    void* staged = nullptr;  // pinned copy of the IQ
    size_t staged_cls = 0;
    bool direct = false;     // `out` is pinned memory of the server's: the IQ is copied there by DMA
    // A request split into pieces (line ranges) owns them and is done when the
    // last one is; a piece writes lines [line0, line0 + lines) of its parent's IQ.
    Request* parent = nullptr;
    int line0 = 0;
    std::vector<std::unique_ptr<Request>> pieces;
    int pieces_left = 0;
};

struct Shape {
    int lines = 0, samples = 0;
    size_t slot_bytes = 0;
};

// One frame context: a batch, its device output, and per request (in the order
// their IQ is copied to host memory, smallest first) the event that marks that
// request's IQ in host memory, so a small request does not wait for the copies
// of larger ones.
struct Frame {
    bcsim_batch batch;
    DevBuf d_out;
    size_t out_accounted = 0;  // bytes of d_out counted in the server's `used`
    std::vector<hipEvent_t> events;  // reused; events[k] follows the copy of copy_order[k]
    hipEvent_t computed = nullptr;   // after the frame's kernels, before its copies
    bool has_piece = false;
    std::vector<Request*> reqs;      // in batch (job) order
    std::vector<Request*> copy_order;
    bool busy = false;
    ~Frame() {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (computed) (void)hipEventDestroy(computed);
    }
};

}  // namespace
}  // namespace bcsim_hip

struct bcsim_server {
    int device = -1;
    size_t budget = 0;       // device bytes for phantoms, frame buffers and outputs
    size_t used = 0;         // accounted device bytes
    size_t resident = 0;     // phantom datasets and LUTs (part of `used`)
    int max_jobs = 64;       // requests per batched frame
    bool autotune = true;
    int piece_lines = 1024;  // requests of more lines are served in pieces of this many (0: never)
    static constexpr int kMaxFrameContexts = 4;
    int contexts = 2;        // frames in flight

    std::mutex mu;
    std::condition_variable cv_sched, cv_comp, cv_done;
    bool stop = false;
    bool sched_exited = false;
    uint64_t next_id = 1;
    uint64_t tick = 0;
    uint64_t completions = 0;

    PinnedPool pinned;
    std::map<const char*, size_t> host_blocks;  // pinned caller buffers (bcsim_server_alloc_host)
    std::vector<std::unique_ptr<bcsim_sim>> phantoms;  // snapshots (views) of the callers' simulators
    std::set<const void*> resident_sets;               // datasets and LUTs already accounted
    std::vector<Geometry> geometries;
    std::deque<Geometry> piece_geoms;  // internal geometries of pieces: id -(k + 1) is piece_geoms[k]
    std::map<std::pair<int, int>, Shape> shapes;
    std::vector<std::unique_ptr<Slot>> slots;
    std::deque<Request*> pending;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    Frame frames[kMaxFrameContexts];
    std::deque<Frame*> inflight;
    Frame* last_piece_frame = nullptr;  // the last frame launched with a piece in it

    // statistics (completion order)
    std::vector<double> lat_ms, queue_ms, copy_ms, frame_jobs, frame_lines, launch_ms;
    double lines_done = 0;
    double t_first_submit = -1, t_last_done = -1;

    std::thread scheduler, completer;

    ~bcsim_server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv_sched.notify_all();
        cv_comp.notify_all();
        if (scheduler.joinable()) scheduler.join();
        if (completer.joinable()) completer.join();
        for (auto& kv : requests) pinned.give(kv.second->staged, kv.second->staged_cls);
        for (auto& kv : host_blocks) (void)hipHostFree(const_cast<char*>(kv.first));
    }

    void start() {
        scheduler = std::thread([this] { scheduler_loop(); });
        completer = std::thread([this] { completer_loop(); });
    }

    // ------------------------------------------------------------ registration
    // Account the datasets and LUT of a phantom not yet counted (shared ones once).
    void account_resident(const bcsim_sim& s) {
        std::set<const void*> added;
        size_t add = 0;
        auto count = [&](const void* key, size_t bytes) {
            if (!resident_sets.count(key) && added.insert(key).second) add += bytes;
        };
        if (s.lut) count(s.lut.get(), s.lut->bytes);
        for (auto& f : s.fixed) count(f.get(), f->pts.bytes + f->boxes.bytes);
        for (auto& sp : s.splines)
            count(sp.get(), sp->cx.bytes + sp->cy.bytes + sp->cz.bytes + sp->amp.bytes + sp->boxes.bytes +
                                sp->iboxes.bytes);
        if (used + add > budget) fail(BCSIM_ERR_OUT_OF_MEMORY, "phantom datasets exceed the server's memory budget");
        resident_sets.insert(added.begin(), added.end());
        resident += add;
        used += add;
    }

    int add_phantom(bcsim_sim* src) {
        need(src);
        if (!src->have_exc) fail(BCSIM_ERR_INVALID_STATE, "phantom has no excitation");
        if (src->profile < 0) fail(BCSIM_ERR_INVALID_STATE, "phantom has no beam profile");
        std::lock_guard<std::mutex> g(mu);
        if (device < 0) device = src->device;
        if (src->device != device) fail(BCSIM_ERR_INVALID_ARGUMENT, "all phantoms of a server must use one device");
        if (budget == 0) {
            hip_check(hipSetDevice(device), "hipSetDevice");
            size_t free_b = 0, total_b = 0;
            hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
            budget = (size_t)(0.9 * (double)free_b);
        }
        auto snap = std::make_unique<bcsim_sim>();
        snap->configure_view_of(*src);
        account_resident(*snap);
        phantoms.push_back(std::move(snap));
        return (int)phantoms.size() - 1;
    }

    int add_geometry(int L, const float* o, const float* d, const float* lat, float length) {
        if (L <= 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "a geometry needs at least one line");
        if (!o || !d || !lat) fail(BCSIM_ERR_INVALID_ARGUMENT, "null geometry array");
        if (!(length > 0)) fail(BCSIM_ERR_INVALID_ARGUMENT, "line length must be positive");
        Geometry gm;
        gm.o.assign(o, o + 3 * (size_t)L);
        gm.d.assign(d, d + 3 * (size_t)L);
        gm.lat.assign(lat, lat + 3 * (size_t)L);
        gm.t.assign(L, 0.0f);
        gm.length = length;
        std::lock_guard<std::mutex> g(mu);
        geometries.push_back(std::move(gm));
        return (int)geometries.size() - 1;
    }

    // A caller's geometry (id >= 0) or a piece's (id < 0).
    const Geometry& geom(int id) const { return id >= 0 ? geometries[id] : piece_geoms[-id - 1]; }

    // The pieces of geometry gi for the current piece size, made on first use (mu held).
    const std::vector<int>& pieces_of(int gi) {
        if (geometries[gi].piece_size != piece_lines) {
            const Geometry src = geometries[gi];
            std::vector<int> ids;
            for (int a = 0; a < src.lines(); a += piece_lines) {
                const int n = std::min(piece_lines, src.lines() - a);
                Geometry pg;
                pg.o.assign(src.o.begin() + 3 * (size_t)a, src.o.begin() + 3 * (size_t)(a + n));
                pg.d.assign(src.d.begin() + 3 * (size_t)a, src.d.begin() + 3 * (size_t)(a + n));
                pg.lat.assign(src.lat.begin() + 3 * (size_t)a, src.lat.begin() + 3 * (size_t)(a + n));
                pg.t.assign(n, 0.0f);
                pg.length = src.length;
                piece_geoms.push_back(std::move(pg));
                ids.push_back(-(int)piece_geoms.size());
            }
            geometries[gi].pieces = std::move(ids);
            geometries[gi].piece_size = piece_lines;
        }
        return geometries[gi].pieces;
    }

    // Simulator for (phantom, geometry) with its frame shape known (nothing allocated).
    std::unique_ptr<bcsim_sim> make_sim(int p, int gi) {
        auto s = std::make_unique<bcsim_sim>();
        s->configure_view_of(*phantoms[p]);
        const Geometry& gm = geom(gi);
        s->set_scan_sequence(gm.lines(), gm.o.data(), gm.d.data(), gm.lat.data(), gm.t.data(), gm.length, true);
        s->update_shape();
        return s;
    }

    // Output shape and frame-buffer estimate of a pair (mu held); gi may be a piece's.
    const Shape& shape_of(int p, int gi) {
        if (p < 0 || p >= (int)phantoms.size()) fail(BCSIM_ERR_INVALID_ARGUMENT, "unknown phantom");
        if (gi >= (int)geometries.size() || gi < -(int)piece_geoms.size())
            fail(BCSIM_ERR_INVALID_ARGUMENT, "unknown geometry");
        auto key = std::make_pair(p, gi);
        auto it = shapes.find(key);
        if (it != shapes.end()) return it->second;
        auto s = make_sim(p, gi);
        Shape sh;
        sh.lines = (int)s->lines.size();
        sh.samples = s->n_out;
        sh.slot_bytes = s->frame_bytes_estimate();
        return shapes[key] = sh;
    }

    // shape_of for a caller's geometry id (mu held).
    const Shape& user_shape(int p, int gi) {
        if (gi < 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "unknown geometry");
        return shape_of(p, gi);
    }

    // ------------------------------------------------------------ requests
    uint64_t submit(int p, int gi, float t, float* out, size_t capacity) {
        if (!out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null output buffer");
        if (!std::isfinite(t)) fail(BCSIM_ERR_INVALID_ARGUMENT, "time must be finite");
        std::lock_guard<std::mutex> g(mu);
        if (stop) fail(BCSIM_ERR_INVALID_STATE, "server is shutting down");
        const Shape& sh = user_shape(p, gi);
        if (capacity < (size_t)sh.lines * sh.samples) fail(BCSIM_ERR_CAPACITY, "output capacity too small");
        auto r = std::make_unique<Request>();
        r->id = next_id++;
        r->phantom = p;
        r->geometry = gi;
        r->t = t;
        r->out = out;
        r->capacity = capacity;
        r->lines = sh.lines;
        r->samples = sh.samples;
        r->direct = in_host_block(out, (size_t)sh.lines * sh.samples * sizeof(float2));
        r->t_submit = now_ms();
        if (t_first_submit < 0) t_first_submit = r->t_submit;
        if (piece_lines > 0 && sh.lines > piece_lines) {
            // served in pieces, each a job of a later frame than the one before,
            // so the first pieces' IQ is copied while the next ones compute
            int line0 = 0;
            for (int pg : pieces_of(gi)) {
                const Shape& ps = shape_of(p, pg);
                if (ps.samples != sh.samples) fail(BCSIM_ERR_INTERNAL, "piece and request differ in samples per line");
                auto q = std::make_unique<Request>();
                q->phantom = p;
                q->geometry = pg;
                q->t = t;
                q->lines = ps.lines;
                q->samples = ps.samples;
                q->parent = r.get();
                q->line0 = line0;
                q->t_submit = r->t_submit;
                line0 += ps.lines;
                r->pieces.push_back(std::move(q));
            }
            r->pieces_left = (int)r->pieces.size();
            for (auto& q : r->pieces) pending.push_back(q.get());
        } else {
            pending.push_back(r.get());
        }
        const uint64_t id = r->id;
        requests[id] = std::move(r);
        cv_sched.notify_one();
        return id;
    }

    // Wait for a request; its record is released once it has been reported.
    void wait(uint64_t id, int timeout_ms, int* nl, int* ns, double* latency) {
        std::unique_lock<std::mutex> lk(mu);
        auto it = requests.find(id);
        if (it == requests.end()) fail(BCSIM_ERR_INVALID_ARGUMENT, "unknown request id");
        Request* r = it->second.get();
        auto done = [&] { return r->done; };
        if (timeout_ms < 0) cv_done.wait(lk, done);
        else if (!cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
            fail(BCSIM_ERR_INVALID_STATE, "request not finished within the timeout");
        if (nl) *nl = r->lines;
        if (ns) *ns = r->samples;
        const int status = r->status;
        const std::string err = r->error;
        std::unique_ptr<Request> own = std::move(it->second);
        requests.erase(it);
        if (status != BCSIM_OK) {
            pinned.give(r->staged, r->staged_cls);
            fail(status, err);
        }
        // copy out in this thread, without the lock (nothing to copy when the DMA
        // wrote the caller's pinned buffer)
        lk.unlock();
        const double c0 = now_ms();
        if (!r->direct) parallel_copy(r->out, r->staged, (size_t)r->lines * r->samples * sizeof(float2));
        const double copy = now_ms() - c0;
        pinned.give(r->staged, r->staged_cls);
        const double lat = r->t_done - r->t_submit + copy;  // as seen by a caller waiting from submit
        if (latency) *latency = lat;
        lk.lock();
        lat_ms.push_back(lat);
        copy_ms.push_back(copy);
    }

    // ------------------------------------------------------------ scheduling (mu held)
    Frame* free_frame() {
        for (int k = 0; k < contexts; k++)
            if (!frames[k].busy) return &frames[k];
        return nullptr;
    }

    Slot* take_idle(int p, int gi) {
        for (auto& s : slots)
            if (!s->busy && s->phantom == p && s->geometry == gi) return s.get();
        return nullptr;
    }

    // Free idle slots, least recently used first, until `need` more bytes fit.
    bool make_room(size_t need, const Slot* keep) {
        while (used + need > budget) {
            Slot* lru = nullptr;
            for (auto& s : slots)
                if (!s->busy && s.get() != keep && (!lru || s->last_used < lru->last_used)) lru = s.get();
            if (!lru) return false;
            used -= lru->bytes;
            for (size_t i = 0; i < slots.size(); i++)
                if (slots[i].get() == lru) {
                    slots.erase(slots.begin() + i);
                    break;
                }
        }
        return true;
    }

    // Is [p, p + bytes) inside one pinned block of bcsim_server_alloc_host? (mu held)
    bool in_host_block(const void* p, size_t bytes) const {
        const char* c = static_cast<const char*>(p);
        auto it = host_blocks.upper_bound(c);
        if (it == host_blocks.begin()) return false;
        --it;
        return c >= it->first && c + bytes <= it->first + it->second;
    }

    void* alloc_host(size_t bytes) {
        if (bytes == 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "zero-byte host buffer");
        hip_check(hipSetDevice(device < 0 ? 0 : device), "hipSetDevice");
        void* p = nullptr;
        hip_check(hipHostMalloc(&p, bytes, 0), "hipHostMalloc");
        std::lock_guard<std::mutex> g(mu);
        host_blocks[static_cast<const char*>(p)] = bytes;
        return p;
    }

    void free_host(void* p) {
        {
            std::lock_guard<std::mutex> g(mu);
            auto it = host_blocks.find(static_cast<const char*>(p));
            if (it == host_blocks.end()) fail(BCSIM_ERR_INVALID_ARGUMENT, "not a host buffer of this server");
            const char* lo = it->first;
            const char* hi = lo + it->second;
            for (auto& kv : requests) {
                const char* o = reinterpret_cast<const char*>(kv.second->out);
                if (o >= lo && o < hi)
                    fail(BCSIM_ERR_INVALID_STATE, "host buffer still holds the output of an unreported request");
            }
            host_blocks.erase(it);
        }
        (void)hipHostFree(p);
    }

    // Returns the request that this completes for its caller: r, its parent when r
    // is the parent's last piece, or null.
    Request* finish(Request* r, int status, const std::string& err, double t) {
        r->status = status;
        r->error = err;
        r->done = true;
        r->t_done = t;
        if (r->slot) {
            r->slot->busy = false;
            r->slot->last_used = ++tick;
            r->slot = nullptr;
        }
        Request* P = r->parent;
        if (!P) return r;
        if (status != BCSIM_OK && P->status == BCSIM_OK) {
            P->status = status;
            P->error = err;
        }
        if (--P->pieces_left > 0) return nullptr;
        P->done = true;
        P->t_done = t;
        return P;
    }

    // The worker threads start with the server, before any phantom names the
    // device: launch() selects it for every frame, and the completer once it has a
    // frame (both read `device` after mu has ordered them behind add_phantom).
    void scheduler_loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_sched.wait(lk, [&] { return stop || (!pending.empty() && free_frame()); });
            if (stop) break;
            Frame* f = free_frame();
            std::vector<Request*> admitted;
            std::set<const Request*> split;  // requests with a piece in this frame: one piece each
            size_t out_elems = 0;
            for (auto it = pending.begin(); it != pending.end() && (int)admitted.size() < max_jobs;) {
                Request* r = *it;
                if (r->parent && split.count(r->parent)) {
                    ++it;  // a later frame takes it
                    continue;
                }
                const Shape& sh = shapes[std::make_pair(r->phantom, r->geometry)];
                Slot* s = take_idle(r->phantom, r->geometry);
                const size_t slot_need = s ? 0 : sh.slot_bytes;
                const size_t out_bytes = (out_elems + (size_t)sh.lines * sh.samples) * sizeof(float2);
                const size_t out_need = out_bytes > f->d_out.bytes ? out_bytes - f->d_out.bytes : 0;
                if (!make_room(slot_need + out_need, s)) {
                    if (admitted.empty() && inflight.empty()) {
                        // nothing will free memory: this request can never run
                        it = pending.erase(it);
                        finish(r, BCSIM_ERR_OUT_OF_MEMORY, "request exceeds the server's memory budget", now_ms());
                        cv_done.notify_all();
                        continue;
                    }
                    break;
                }
                if (!s) {
                    auto slot = std::make_unique<Slot>();
                    try {
                        slot->sim = make_sim(r->phantom, r->geometry);
                    } catch (const BcsimError& e) {
                        it = pending.erase(it);
                        finish(r, e.code, e.what(), now_ms());
                        cv_done.notify_all();
                        continue;
                    }
                    slot->phantom = r->phantom;
                    slot->geometry = r->geometry;
                    slot->bytes = sh.slot_bytes;
                    used += slot->bytes;
                    s = slot.get();
                    slots.push_back(std::move(slot));
                }
                s->busy = true;
                r->slot = s;
                it = pending.erase(it);
                admitted.push_back(r);
                if (r->parent) split.insert(r->parent);
                out_elems += (size_t)sh.lines * sh.samples;
            }
            if (admitted.empty()) {
                if (pending.empty()) continue;
                const uint64_t seen = completions;  // wait until a frame finishes and frees memory
                cv_sched.wait(lk, [&] { return stop || completions != seen; });
                continue;
            }
            f->busy = true;
            lk.unlock();
            std::string err;
            int code = BCSIM_OK;
            const double l0 = now_ms();
            try {
                launch(f, admitted);
            } catch (const BcsimError& e) {
                code = e.code;
                err = e.what();
            } catch (const std::exception& e) {
                code = BCSIM_ERR_INTERNAL;
                err = e.what();
            }
            const double l1 = now_ms();
            lk.lock();
            launch_ms.push_back(l1 - l0);
            if (code == BCSIM_OK && f->reqs.empty()) {  // every admitted request failed in preparation
                f->busy = false;
                continue;
            }
            if (code != BCSIM_OK) {
                // copies the launch enqueued before it failed may still write the buffers
                lk.unlock();
                (void)hipStreamSynchronize(f->batch.stream);
                lk.lock();
                const double t = now_ms();
                for (Request* r : f->reqs) finish(r, code, err, t);
                f->reqs.clear();
                f->busy = false;
                cv_done.notify_all();
                continue;
            }
            // account the frame buffers as allocated
            for (Request* r : f->reqs) {
                const size_t actual = r->slot->sim->frame_bytes();
                used = used - r->slot->bytes + actual;
                r->slot->bytes = actual;
            }
            inflight.push_back(f);
            cv_comp.notify_one();
        }
        // fail what never launched
        const double t = now_ms();
        for (Request* r : pending) finish(r, BCSIM_ERR_INVALID_STATE, "server shut down before the request ran", t);
        pending.clear();
        sched_exited = true;
        cv_comp.notify_all();
        cv_done.notify_all();
    }

    // Prepare the admitted requests (outside mu; their slots are ours) and launch
    // them as one batched frame followed by the copy to pinned memory.
    void launch(Frame* f, std::vector<Request*>& admitted) {
        hip_check(hipSetDevice(device), "hipSetDevice");
        f->reqs.clear();
        for (Request* r : admitted) {
            bcsim_sim* s = r->slot->sim.get();
            try {
                s->set_frame_time(r->t);
                s->prepare();
                f->reqs.push_back(r);
            } catch (const BcsimError& e) {
                std::lock_guard<std::mutex> g(mu);
                finish(r, e.code, e.what(), now_ms());
                cv_done.notify_all();
            }
        }
        if (f->reqs.empty()) return;
        f->has_piece = false;
        for (Request* r : f->reqs) {
            r->t_launch = now_ms();
            if (Request* P = r->parent) {
                f->has_piece = true;
                if (P->t_launch == 0) P->t_launch = r->t_launch;
            }
        }
        bcsim_batch& b = f->batch;
        b.use_graph = false;  // the mix changes from frame to frame
        b.autotune = autotune;
        b.sims.clear();
        b.device = -1;
        b.invalidate();
        for (Request* r : f->reqs) b.add(r->slot->sim.get());
        b.build();
        const size_t total = b.total_out;
        if (f->d_out.bytes < total * sizeof(float2)) {
            // by at least half again: a hipFree waits for every stream of the device
            hip_check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
            f->d_out.ensure(std::max(total * sizeof(float2), f->d_out.bytes + f->d_out.bytes / 2));
        }
        for (Request* r : f->reqs) {
            Request* o = r->parent ? r->parent : r;  // a piece is copied into its parent's buffer
            if (!o->staged && !o->direct)
                o->staged = pinned.take((size_t)o->lines * o->samples * sizeof(float2), o->staged_cls);
        }
        while (f->events.size() < f->reqs.size()) {
            hipEvent_t e = nullptr;
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            f->events.push_back(e);
        }
        if (!f->computed) hip_check(hipEventCreateWithFlags(&f->computed, hipEventDisableTiming), "hipEventCreate");
        // A frame with a piece computes after the previous such frame's kernels, so
        // the pieces of a request finish one after another and each one's copy
        // overlaps the next one's kernels (rather than all finishing together).
        if (f->has_piece) {
            if (last_piece_frame && last_piece_frame != f)
                hip_check(hipStreamWaitEvent(b.stream, last_piece_frame->computed, 0), "hipStreamWaitEvent");
            last_piece_frame = f;
        }
        b.simulate_device(f->d_out.p, total);
        hip_check(hipEventRecord(f->computed, b.stream), "hipEventRecord");
        // IQ to host memory, smallest first, an event after each request's copy
        std::vector<size_t> idx(f->reqs.size());
        for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t c) {
            return (size_t)f->reqs[a]->lines * f->reqs[a]->samples < (size_t)f->reqs[c]->lines * f->reqs[c]->samples;
        });
        f->copy_order.clear();
        for (size_t k = 0; k < idx.size(); k++) {
            Request* r = f->reqs[idx[k]];
            const float2* src = f->d_out.as<float2>() + b.plan[idx[k]].out_offset;
            const Request* o = r->parent ? r->parent : r;
            float2* dst = static_cast<float2*>(o->direct ? (void*)o->out : o->staged) + (size_t)r->line0 * r->samples;
            hip_check(hipMemcpyAsync(dst, src, (size_t)r->lines * r->samples * sizeof(float2), hipMemcpyDeviceToHost,
                                     b.stream),
                      "hipMemcpyAsync");
            hip_check(hipEventRecord(f->events[k], b.stream), "hipEventRecord");
            f->copy_order.push_back(r);
        }
        std::lock_guard<std::mutex> g(mu);
        used = used - f->out_accounted + f->d_out.bytes;
        f->out_accounted = f->d_out.bytes;
        frame_jobs.push_back((double)f->reqs.size());
        double nl = 0;
        for (Request* r : f->reqs) nl += r->lines;
        frame_lines.push_back(nl);
    }

    void completer_loop() {
        std::unique_lock<std::mutex> lk(mu);
        int current_device = -1;
        for (;;) {
            cv_comp.wait(lk, [&] { return !inflight.empty() || sched_exited; });
            if (inflight.empty()) break;
            Frame* f = inflight.front();
            if (current_device != device) {  // mu held: set by add_phantom before any frame
                (void)hipSetDevice(device);  // a failure shows in the event waits below
                current_device = device;
            }
            // requests complete one by one, as their copies land
            for (size_t k = 0; k < f->copy_order.size(); k++) {
                Request* r = f->copy_order[k];
                lk.unlock();
                const hipError_t e = hipEventSynchronize(f->events[k]);
                const double t = now_ms();
                lk.lock();
                if (e == hipSuccess) {
                    const Request* whole = finish(r, BCSIM_OK, "", t);
                    if (whole) queue_ms.push_back(whole->t_launch - whole->t_submit);
                    lines_done += r->lines;
                    t_last_done = t;
                } else {
                    finish(r, BCSIM_ERR_DEVICE, std::string("frame failed: ") + hipGetErrorString(e), t);
                }
                cv_done.notify_all();
            }
            f->reqs.clear();
            f->copy_order.clear();
            f->busy = false;
            inflight.pop_front();
            completions++;
            cv_sched.notify_all();
            cv_done.notify_all();
        }
    }

    void get_stats(const std::string& key, std::vector<double>& out) {
        std::lock_guard<std::mutex> g(mu);
        if (key == "latency_ms") out = lat_ms;
        else if (key == "queue_ms") out = queue_ms;
        else if (key == "copy_ms") out = copy_ms;
        else if (key == "frame_jobs") out = frame_jobs;
        else if (key == "frame_lines") out = frame_lines;
        else if (key == "launch_ms") out = launch_ms;
        else if (key == "lines_completed") out = {lines_done};
        else if (key == "busy_span_ms") out = {t_first_submit >= 0 && t_last_done >= 0 ? t_last_done - t_first_submit : 0.0};
        else if (key == "budget_bytes") out = {(double)budget};
        else if (key == "used_bytes") out = {(double)used};
        else if (key == "resident_bytes") out = {(double)resident};
        else if (key == "idle_slots") {
            double n = 0;
            for (auto& s : slots) n += !s->busy;
            out = {n};
        } else if (key == "pending") out = {(double)pending.size()};
        else fail(BCSIM_ERR_INVALID_ARGUMENT, "invalid statistics key: " + key);
    }

};

extern "C" {

BCSIM_API bcsim_status bcsim_server_create(bcsim_server** out) {
    return guarded([&] {
        if (!out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null output handle");
        auto s = std::make_unique<bcsim_server>();
        s->start();
        *out = s.release();
    });
}

BCSIM_API bcsim_status bcsim_server_destroy(bcsim_server* s) {
    return guarded([&] { delete s; });
}

BCSIM_API bcsim_status bcsim_server_set_parameter(bcsim_server* s, const char* key, const char* value) {
    return guarded([&] {
        if (!s || !key || !value) fail(BCSIM_ERR_INVALID_ARGUMENT, "null argument");
        const std::string k = key;
        std::lock_guard<std::mutex> g(s->mu);
        if (k == "memory_budget_bytes") {
            const long long v = std::stoll(value);
            if (v <= 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "memory budget must be positive");
            if ((size_t)v < s->used) fail(BCSIM_ERR_INVALID_ARGUMENT, "memory budget below the bytes already in use");
            s->budget = (size_t)v;
        } else if (k == "max_jobs_per_frame") {
            const int v = parse_int(value);
            if (v < 1) fail(BCSIM_ERR_INVALID_ARGUMENT, "max_jobs_per_frame must be >= 1");
            s->max_jobs = v;
        } else if (k == "autotune") {
            s->autotune = parse_bool(value);
        } else if (k == "piece_lines") {
            const int v = parse_int(value);
            if (v < 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "piece_lines must be >= 0");
            s->piece_lines = v;
        } else if (k == "frame_contexts") {
            const int v = parse_int(value);
            if (v < 1 || v > bcsim_server::kMaxFrameContexts)
                fail(BCSIM_ERR_INVALID_ARGUMENT, "frame_contexts must be in [1, 4]");
            s->contexts = v;
        } else if (k == "reset_stats") {
            s->lat_ms.clear();
            s->queue_ms.clear();
            s->copy_ms.clear();
            s->frame_jobs.clear();
            s->frame_lines.clear();
            s->launch_ms.clear();
            s->lines_done = 0;
            s->t_first_submit = s->t_last_done = -1;
        } else {
            fail(BCSIM_ERR_INVALID_ARGUMENT, "illegal parameter name: '" + k + "'");
        }
        s->cv_sched.notify_all();
    });
}

BCSIM_API bcsim_status bcsim_server_add_phantom(bcsim_server* s, bcsim_sim* phantom, int* phantom_id) {
    return guarded([&] {
        if (!s) fail(BCSIM_ERR_INVALID_ARGUMENT, "null server handle");
        const int id = s->add_phantom(phantom);
        if (phantom_id) *phantom_id = id;
    });
}

BCSIM_API bcsim_status bcsim_server_add_geometry(bcsim_server* s, int num_lines, const float* origins,
                                                 const float* directions, const float* lateral_dirs,
                                                 float line_length, int* geometry_id) {
    return guarded([&] {
        if (!s) fail(BCSIM_ERR_INVALID_ARGUMENT, "null server handle");
        const int id = s->add_geometry(num_lines, origins, directions, lateral_dirs, line_length);
        if (geometry_id) *geometry_id = id;
    });
}

BCSIM_API bcsim_status bcsim_server_get_output_shape(bcsim_server* s, int phantom_id, int geometry_id,
                                                     int* num_lines, int* num_samples) {
    return guarded([&] {
        if (!s) fail(BCSIM_ERR_INVALID_ARGUMENT, "null server handle");
        std::lock_guard<std::mutex> g(s->mu);
        const Shape& sh = s->user_shape(phantom_id, geometry_id);
        if (num_lines) *num_lines = sh.lines;
        if (num_samples) *num_samples = sh.samples;
    });
}

BCSIM_API bcsim_status bcsim_server_submit(bcsim_server* s, int phantom_id, int geometry_id, float t, float* out,
                                           size_t capacity, uint64_t* request_id) {
    return guarded([&] {
        if (!s) fail(BCSIM_ERR_INVALID_ARGUMENT, "null server handle");
        const uint64_t id = s->submit(phantom_id, geometry_id, t, out, capacity);
        if (request_id) *request_id = id;
    });
}

BCSIM_API bcsim_status bcsim_server_wait(bcsim_server* s, uint64_t request_id, int timeout_ms, int* num_lines,
                                         int* num_samples, double* latency_ms) {
    return guarded([&] {
        if (!s) fail(BCSIM_ERR_INVALID_ARGUMENT, "null server handle");
        s->wait(request_id, timeout_ms, num_lines, num_samples, latency_ms);
    });
}

BCSIM_API bcsim_status bcsim_server_get_stats(bcsim_server* s, const char* key, double* out, size_t capacity,
                                              size_t* count) {
    return guarded([&] {
        if (!s || !key) fail(BCSIM_ERR_INVALID_ARGUMENT, "null argument");
        std::vector<double> v;
        s->get_stats(key, v);
        if (count) *count = v.size();
        const size_t n = std::min(capacity, v.size());
        if (out && n) std::memcpy(out, v.data(), n * sizeof(double));
    });
}

BCSIM_API bcsim_status bcsim_server_alloc_host(bcsim_server* s, size_t bytes, void** ptr) {
    return guarded([&] {
        if (!s || !ptr) fail(BCSIM_ERR_INVALID_ARGUMENT, "null argument");
        *ptr = s->alloc_host(bytes);
    });
}

BCSIM_API bcsim_status bcsim_server_free_host(bcsim_server* s, void* ptr) {
    return guarded([&] {
        if (!s || !ptr) fail(BCSIM_ERR_INVALID_ARGUMENT, "null argument");
        s->free_host(ptr);
    });
}

}  // extern "C"
