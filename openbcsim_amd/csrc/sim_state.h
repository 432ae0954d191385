// Host-side state of one simulator handle (the "gpu" algorithm), shared by the
// single-simulator C-ABI (sim_host.cpp) and the batched / serving paths
// (batch_host.cpp).  See sim_host.cpp for the role of each part.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bcsim_c.h"
#include "autotune.h"
#include "host_copy.h"
#include "sim_device.h"


namespace bcsim_hip {

std::string& last_error();  // this thread's message for bcsim_last_error()


struct BcsimError : std::runtime_error {
    int code;
    BcsimError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw BcsimError(code, msg); }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? BCSIM_ERR_OUT_OF_MEMORY
                                                                                      : BCSIM_ERR_DEVICE;
        fail(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// Device buffer owned by a simulator.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n) {
        if (n <= bytes && p) return;
        release();
        if (n == 0) return;
        hip_check(hipMalloc(&p, n), "hipMalloc");
        bytes = n;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// Datasets are stored in Morton (z-curve) order of position, in chunks of
// kChunk scatterers with one bounding box each, so the tile kernel can skip
// whole chunks that lie outside a beam.  The order of the scatterers does not
// change the frame beyond float summation order.  A dataset is immutable once
// built and shared (shared_ptr) by every simulator that projects it: the
// serving path keeps one resident copy of a phantom for many frame requests.
// Orders in which the tile kernel visits a dataset's chunks, one per scan
// origin: by the distance of the chunk's box centre from a common origin, or by
// its depth along the middle line when the lines do not share one.  Every
// workgroup then sweeps the dataset from near to far at about the same pace, so
// the workgroups sharing an L2 read nearby radial slices of the beam-profile
// table at any one time.  The order never changes the sums.  Orders live as long as the dataset (at
// most kMaxOrders origins; further origins visit the chunks in index order).  An order from a
// common origin is followed in the same buffer by two bounds per position p of the order, the
// largest far distance (centre distance + box half-diagonal) over positions <= p and the smallest
// near distance (centre distance - half-diagonal) over positions >= p; both ascend with p, so a
// radial window of the tile kernel finds by two binary searches the positions whose chunks can
// reach its distances.  A chunk with a large box (where the z-curve jumps) widens only the
// windows near its own distance.
struct ChunkOrders {
    using Key = std::array<float, 7>;  // origin, direction (depth mode), shared-origin flag
    static constexpr size_t kMaxOrders = 16;
    std::vector<float4> centres;       // box centres; w = 1 marks an empty chunk
    std::vector<float> half_diag;      // box half-diagonals
    std::mutex mu;
    std::vector<std::pair<Key, std::unique_ptr<DevBuf>>> orders;

    void set_boxes(const std::vector<ChunkBox>& boxes) {
        centres.resize(boxes.size());
        half_diag.resize(boxes.size());
        for (size_t c = 0; c < boxes.size(); c++) {
            const ChunkBox& b = boxes[c];
            centres[c] = make_float4(b.cx, b.cy, b.cz, b.hx >= 0.0f ? 0.0f : 1.0f);
            half_diag[c] = b.hx >= 0.0f ? (float)(std::sqrt((double)b.hx * b.hx + (double)b.hy * b.hy +
                                                            (double)b.hz * b.hz) * (1.0 + 1e-6))
                                        : 0.0f;
        }
    }

    // The bounds that follow a common-origin order of n chunks: far[n] then near[n], or null.
    static const float* distances(const Key& key, const int* order, int n) {
        return key[6] != 0.0f && order ? reinterpret_cast<const float*>(order + n) : nullptr;
    }

    // Device array of chunk indices (a permutation of [0, nchunks)), or null.
    // create = false: only an order built earlier (no allocation).
    const int* get(const Key& key, bool create = true) {
        std::lock_guard<std::mutex> g(mu);
        for (auto& e : orders)
            if (e.first == key) return e.second->as<int>();
        if (!create || orders.size() >= kMaxOrders || centres.empty()) return nullptr;
        std::vector<std::pair<float, int>> k(centres.size());
        for (size_t c = 0; c < centres.size(); c++) {
            const float4 q = centres[c];
            const float dx = q.x - key[0], dy = q.y - key[1], dz = q.z - key[2];
            float v = key[6] != 0.0f ? dx * dx + dy * dy + dz * dz : dx * key[3] + dy * key[4] + dz * key[5];
            if (q.w != 0.0f || !(v == v)) v = INFINITY;  // empty chunks (and NaN keys) last
            k[c] = {v, (int)c};
        }
        std::sort(k.begin(), k.end());
        const bool shared = key[6] != 0.0f;
        const size_t n = k.size();
        std::vector<int> order(n * (shared ? 3 : 1));
        for (size_t c = 0; c < n; c++) order[c] = k[c].second;
        if (shared) {
            // running max of far distances from the front, running min of near ones from the back
            float far = -INFINITY, near = INFINITY;
            for (size_t c = 0; c < n; c++) {
                const float d = std::sqrt(k[c].first), h = half_diag[k[c].second];
                far = std::max(far, d + h);  // infinite from the first empty chunk on
                std::memcpy(&order[n + c], &far, sizeof far);
            }
            for (size_t c = n; c-- > 0;) {
                const float d = std::sqrt(k[c].first), h = half_diag[k[c].second];
                near = std::min(near, d - h);
                std::memcpy(&order[2 * n + c], &near, sizeof near);
            }
        }
        auto buf = std::make_unique<DevBuf>();
        buf->ensure(order.size() * sizeof(int));
        hip_check(hipMemcpy(buf->p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
        orders.emplace_back(key, std::move(buf));
        return orders.back().second->as<int>();
    }
};

struct FixedDataset {
    DevBuf pts;    // float4[n] (x, y, z, amplitude), Morton order
    DevBuf boxes;  // ChunkBox[nchunks]
    size_t n = 0;
    int nchunks = 0;
    double peak_amp = 0.0;  // largest |amplitude| (infinite if any is not finite)
    double typ_amp = 0.0;   // median |amplitude| of the non-zero amplitudes
    mutable ChunkOrders visit;
};

struct SplineDataset {
    int degree = 0;
    int num_cp = 0;
    size_t n = 0;
    int nchunks = 0;
    double peak_amp = 0.0;  // largest |amplitude| (infinite if any is not finite)
    double typ_amp = 0.0;   // median |amplitude| of the non-zero amplitudes
    std::vector<float> knots;
    DevBuf cx, cy, cz, amp;  // [num_cp][n] x3, [n], Morton order of the control-point centroid
    DevBuf boxes;            // ChunkBox[nchunks] over all control points of the chunk
    DevBuf iboxes;           // ChunkBox[nchunks][num_cp - degree]: per knot interval (SplineSet::iboxes)
    // the interval boxes are used when they are tighter than the chunk boxes (mean
    // half-extent below 0.95 of the chunk boxes': scatterers that move far in a frame);
    // otherwise their extra load per chunk and line costs more than they cull
    // (measured on the 200k-spline M-mode phantom with 0.5 mm of jitter: 0.1265 ->
    // 0.1287 ms with them, against 0.1393 -> 0.1361 ms on a phantom contracting by 8 mm)
    bool use_iboxes = false;
    mutable ChunkOrders visit;
};

// Per-simulator buffers of one spline dataset that depend on the scan.
struct SplineScan {
    DevBuf basis;     // SplineLineBasis[num_lines]
    DevBuf rendered;  // float4[n], all-timestamps-equal path (culled with the dataset's boxes,
                      // since a rendered point lies in the box of its control points)
};

// Largest |a[i * stride]| over n values (infinite if one of them is not
// finite) and the median |a| of the non-zero ones (a typical amplitude that a
// few strong scatterers do not move).
inline void amplitude_stats(const float* a, size_t n, size_t stride, double& peak, double& typical) {
    std::vector<float> mag;
    mag.reserve(n);
    float m = 0.0f;
    for (size_t i = 0; i < n; i++) {
        const float v = std::fabs(a[i * stride]);
        if (!(v <= 3.4e38f)) {
            peak = typical = INFINITY;
            return;
        }
        m = std::max(m, v);
        if (v > 0.0f) mag.push_back(v);
    }
    peak = m;
    typical = 0.0;
    if (!mag.empty()) {
        std::nth_element(mag.begin(), mag.begin() + mag.size() / 2, mag.end());
        typical = mag[mag.size() / 2];
    }
}

// Fixed-point scale of a projection launch (ProjCommon::amp_scale): the largest
// power of two 2^k with peak * 2^k < 2^50 and peak * n * 2^k < 2^61, where peak
// bounds every |w*a| of the launch (to float rounding) and n is its number of
// scatterers: each scaled contribution is below 2^51 (the kernel's conversion
// needs that) and no bin sum can overflow 64 bits.  The grid step 2^-k must also
// stay within 2^-22 of a typical full-weight contribution (`typical`: the median
// non-zero |a| times the profile's peak), else amplitudes span more than the
// fixed point resolves (peak/typical beyond ~2^39/n) and the launch keeps float
// bins: amp_scale 0, as for a peak that is not finite.
inline void set_fixed_scale(ProjCommon& c, double peak, double typical, size_t n) {
    c.amp_scale = 0.0f;
    c.inv_amp_scale = 0.0f;
    if (!std::isfinite(peak)) return;
    int k = 0;
    if (peak > 0.0) {
        int e1 = 0, e2 = 0;
        (void)std::frexp(peak, &e1);                                  // peak < 2^e1
        (void)std::frexp(peak * (double)std::max<size_t>(n, 1), &e2);  // peak * n < 2^e2
        k = std::max(-126, std::min({126, 50 - e1, 61 - e2}));
        if (std::ldexp(1.0, -k) > std::ldexp(typical, -22)) return;
    }
    c.amp_scale = std::ldexp(1.0f, k);
    c.inv_amp_scale = std::ldexp(1.0f, -k);
}

// 10 bits of each coordinate interleaved into a 30-bit z-curve key.
inline uint32_t spread10(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

inline bool finite3(const float* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

// Permutation that puts the points pos[i*stride .. +2] in z-curve order over
// their bounding box; non-finite points go last.
inline std::vector<uint32_t> morton_order(const float* pos, size_t n, size_t stride) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; i++) {
        const float* p = pos + i * stride;
        if (!finite3(p)) continue;
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    std::vector<std::pair<uint32_t, uint32_t>> keyed(n);
    for (size_t i = 0; i < n; i++) {
        const float* p = pos + i * stride;
        uint32_t key = 0xffffffffu;
        if (finite3(p)) {
            key = 0;
            for (int a = 0; a < 3; a++) {
                const double span = (double)hi[a] - (double)lo[a];
                double q = span > 0 ? ((double)p[a] - (double)lo[a]) / span * 1024.0 : 0.0;
                const uint32_t qi = (uint32_t)std::min(1023.0, std::max(0.0, q));
                key |= spread10(qi) << a;
            }
        }
        keyed[i] = {key, (uint32_t)i};
    }
    std::sort(keyed.begin(), keyed.end());
    std::vector<uint32_t> perm(n);
    for (size_t i = 0; i < n; i++) perm[i] = keyed[i].second;
    return perm;
}

// Box of a set of points, padded by a relative 1e-6 so that rounding of the
// centre and half extent never shrinks it.
struct BoxAcc {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool any = false;
    void add(const float* p) {
        if (!finite3(p)) return;
        any = true;
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    ChunkBox box() const {
        ChunkBox b{0.f, 0.f, 0.f, 0.f, -1.f, -1.f, -1.f, 0.f};
        if (!any) return b;
        float c[3], h[3];
        for (int a = 0; a < 3; a++) {
            const double cd = 0.5 * ((double)lo[a] + (double)hi[a]);
            c[a] = (float)cd;
            const double hd = std::max((double)hi[a] - (double)c[a], (double)c[a] - (double)lo[a]);
            h[a] = (float)(hd * (1.0 + 1e-6) + 1e-30);
        }
        b = ChunkBox{c[0], c[1], c[2], 0.f, h[0], h[1], h[2], 0.f};
        return b;
    }
};

// Vector3D<float>::dot evaluated left to right (vector3.hpp:86-88).
inline float dot3(const float* a, const float* b) {
    float t = a[0] * b[0];
    t = t + a[1] * b[1];
    t = t + a[2] * b[2];
    return t;
}

// bspline_storve::bsplineBasis<float> (bspline.hpp:45-70): the float recursion
// whose values the reference multiplies with the control points.
inline float special_div(float num, float den) {
    if (std::fabs((double)num) < 1e-6 && std::fabs((double)den) < 1e-6) return 0.0f;
    return num / den;
}

inline float bspline_basis(int j, int p, float x, const std::vector<float>& knots) {
    if (p == 0) return (x >= knots[j] && x < knots[j + 1]) ? 1.0f : 0.0f;
    const float left = special_div((x - knots[j]) * bspline_basis(j, p - 1, x, knots), knots[j + p] - knots[j]);
    const float right = special_div((knots[j + 1 + p] - x) * bspline_basis(j + 1, p - 1, x, knots),
                                    knots[j + 1 + p] - knots[j + 1]);
    return left + right;
}

// bspline_storve::compute_knot_interval (bspline.hpp:75-82)
inline int knot_interval(const std::vector<float>& knots, float t) {
    for (int i = 0; i + 1 < (int)knots.size(); i++)
        if (t >= knots[i] && t < knots[i + 1]) return i;
    fail(BCSIM_ERR_INVALID_ARGUMENT, "compute_knot_interval : could not determine knot interval");
}

inline bool parse_bool(const std::string& v) {
    if (v == "on" || v == "true") return true;
    if (v == "off" || v == "false") return false;
    fail(BCSIM_ERR_INVALID_ARGUMENT, "invalid boolean value");
}

inline int parse_int(const std::string& v) {
    try {
        size_t pos = 0;
        const int x = std::stoi(v, &pos);
        return x;
    } catch (...) {
        fail(BCSIM_ERR_INVALID_ARGUMENT, "invalid integer value: '" + v + "'");
    }
}

// A whole decimal number in [0, 2^32); anything else (signs, trailing text,
// overflow) is an invalid argument.
inline uint32_t parse_u32(const std::string& v) {
    if (v.empty() || v.size() > 10 || v.find_first_not_of("0123456789") != std::string::npos)
        fail(BCSIM_ERR_INVALID_ARGUMENT, "invalid unsigned 32-bit integer: '" + v + "'");
    const unsigned long long x = std::stoull(v);
    if (x > 0xffffffffull) fail(BCSIM_ERR_INVALID_ARGUMENT, "invalid unsigned 32-bit integer: '" + v + "'");
    return (uint32_t)x;
}

inline float parse_float(const std::string& v) {
    try {
        return std::stof(v);
    } catch (...) {
        fail(BCSIM_ERR_INVALID_ARGUMENT, "invalid number: '" + v + "'");
    }
}

}  // namespace bcsim_hip

using namespace bcsim_hip;

// -------------------------------------------------------------------------------------------

struct bcsim_sim {
    // parameters (BaseAlgorithm.cpp:36-46, GpuAlgorithm.cpp:46-60 defaults)
    int verbose = 0;
    float sound_speed = 1540.0f;
    float noise_amplitude = 0.0f;
    uint32_t noise_seed = 0;
    uint32_t noise_counter = 0;
    bool arc = true;
    int decimation = 1;
    bool phase_delay = false;
    int device = 0;
    bool can_change_device = true;
    int num_streams = 2;
    int threads_per_block = 128;
    bool store_details = false;
    bool count_pairs = false;  // projection_counters: the launches count their work (debug data)
    // Beam-profile weights below this fraction of the profile's peak magnitude
    // (1 for the Gaussian, the largest |sample| of a LUT) count as zero, so their
    // pairs are culled; 0 keeps every pair whose weight is not exactly zero.
    float weight_cutoff = kDefaultWeightCutoff;
    uint64_t version = 0;      // bumped by every setter; batches rebuild their plans on change
    bool direct_mode = false;  // projection_mode: "tiled" (culled tile kernel) or "direct"
    // launch shape of the tile kernel set by the user; 0 = tuned or heuristic
    TileTune tune;
    bool autotune = false;  // time candidate launch shapes once per shape class
    // Tile launches of more than one round of workgroups run their blocks
    // heaviest first, by the work each block took in this simulator's previous
    // frame (counted by the launch itself); the order never changes the sums.
    bool tile_order = true;
    bool chunk_order = true;  // visit chunks near to far (ChunkOrders)
    // bin_accumulation: the tile kernel's LDS bins as 64-bit fixed point ("fixed",
    // integer LDS atomics; exact sums of contributions rounded to a grid of about
    // 2^-61 of the launch's largest possible bin), float ("float", ds_add_f32), or
    // "auto" (0): tuned where tuning is on, else fixed for the Gaussian profile and
    // float for a LUT (target_fixed).  Datasets whose amplitudes the grid does not
    // resolve (a non-finite amplitude, or a peak beyond ~2^39/n times the median;
    // set_fixed_scale) use float bins.
    int bin_mode = 0;  // 0 auto, 1 fixed, 2 float (TileTune::bins)
    std::vector<std::unique_ptr<DevBuf>> tile_order_bufs;  // per projection target: [work | order]
    int radix = 0;          // FFT pass radix set by the user (4 or 8); 0 = tuned or default
    std::string last_plan;  // launch shapes of the last frame, for get_parameter("launch_plan")

    hipStream_t stream = nullptr;
    bcsim_log_fn log_fn = nullptr;
    void* log_user = nullptr;

    // excitation
    bool have_exc = false;
    std::vector<float> exc;
    int center = 0;
    float fs = 0.0f;
    float f_demod = 0.0f;

    // scan
    bool have_scan = false;
    std::vector<LineGeom> lines;
    float line_length = 0.0f;
    bool all_equal = false;

    // phantom
    std::vector<std::shared_ptr<const FixedDataset>> fixed;
    std::vector<std::shared_ptr<const SplineDataset>> splines;
    std::vector<std::unique_ptr<SplineScan>> spline_scan;  // parallel to `splines`

    // profile
    int profile = -1;
    float sigma_l = 0.f, sigma_e = 0.f;
    std::shared_ptr<DevBuf> lut;  // float4 quad pairs [n_rad][n_ele][n_lat][2]; replaced, never rewritten
    std::shared_ptr<const std::vector<float>> lut_samples;  // the table as given, for the cutoff's boxes
    float lut_peak = 0.0f;                                   // largest |sample|
    std::shared_ptr<DevBuf> lut_boxes;                       // ProfileParams::slice_box for lut_boxes_cut
    float lut_boxes_cut = -1.0f;
    int lut_nr = 0, lut_nl = 0, lut_ne = 0;
    float lut_r0 = 0, lut_r1 = 0, lut_l0 = 0, lut_l1 = 0, lut_e0 = 0, lut_e1 = 0;

    // derived device tables
    bool dirty_scan = true, dirty_conv = true, dirty_basis = true;
    int num_rf = 0, fft_len = 0, log2_len = 0, n_out = 0;
    int tables_fft = 0, tables_out = 0;  // frame shape the convolution tables were built for
    DevBuf d_lines, d_H, d_demod, d_tw, d_bins, d_iq;
    DevBuf d_stats;  // with store_kernel_details: per projection launch, weight evaluations and LDS additions

    std::map<std::string, std::vector<double>> debug;

    ~bcsim_sim() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (int i = 0; i < 2; i++) {
            if (stage[i]) (void)hipHostFree(stage[i]);
            if (stage_ev[i]) (void)hipEventDestroy(stage_ev[i]);
        }
    }

    void log(int level, const std::string& msg) {
        if (log_fn) log_fn(level, msg.c_str(), log_user);
    }

    void use_device() {
        hip_check(hipSetDevice(device), "hipSetDevice");
        if (!stream) {
            hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
            // set-up of the device: the address of its allocation-free scratch
            void* scratch_addr = nullptr;
            hip_check(noalloc_scratch(device, &scratch_addr), "hipGetSymbolAddress(scratch)");
        }
        can_change_device = false;
    }

    // Wait for frames in flight (simulate_lines_device is asynchronous) before
    // device buffers they read are rewritten, reallocated or freed.
    void quiesce() {
        if (stream) hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    // ---------------------------------------------------------------- parameters
    void set_parameter(const std::string& key, const std::string& value) {
        ++version;
        if (key == "verbose") {
            verbose = parse_int(value);
        } else if (key == "sound_speed") {
            const float c = parse_float(value);
            if (c <= 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "illegal sound speed");
            sound_speed = c;
            dirty_scan = dirty_conv = true;
        } else if (key == "noise_amplitude") {
            noise_amplitude = parse_float(value);
        } else if (key == "noise_seed") {
            noise_seed = parse_u32(value);
            noise_counter = 0;
        } else if (key == "use_arc_projection") {
            arc = parse_bool(value);
        } else if (key == "radial_decimation") {
            const int d = parse_int(value);
            if (d <= 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "illegal radial decimation value");
            decimation = d;
            dirty_conv = true;
        } else if (key == "phase_delay") {
            phase_delay = parse_bool(value);
        } else if (key == "gpu_device") {
            if (!can_change_device) fail(BCSIM_ERR_INVALID_STATE, "cannot change CUDA device now");
            int count = 0;
            hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
            const int d = parse_int(value);
            if (d < 0 || d >= count) fail(BCSIM_ERR_INVALID_ARGUMENT, "illegal device number");
            device = d;
        } else if (key == "cuda_streams") {
            const int n = parse_int(value);
            if (n > 2) fail(BCSIM_ERR_INVALID_ARGUMENT, "number of CUDA streams exceeds MAX_NUM_CUDA_STREAMS");
            if (n <= 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "number of CUDA streams must be more than zero");
            num_streams = n;  // a frame is one stream-ordered launch sequence; kept for compatibility
        } else if (key == "threads_per_block") {
            const int n = parse_int(value);
            if (n <= 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "invalid number of threads per block");
            threads_per_block = n;  // launch shapes come from the tuner; kept for compatibility
        } else if (key == "store_kernel_details") {
            store_details = parse_bool(value);
        } else if (key == "projection_counters") {
            count_pairs = parse_bool(value);
        } else if (key == "weight_cutoff") {
            const float v = parse_float(value);
            if (!(v >= 0.0f && v < 1.0f)) fail(BCSIM_ERR_INVALID_ARGUMENT, "weight_cutoff must lie in [0, 1)");
            weight_cutoff = v;
        } else if (key == "projection_mode") {
            if (value == "tiled") direct_mode = false;
            else if (value == "direct") direct_mode = true;
            else fail(BCSIM_ERR_INVALID_ARGUMENT, "projection_mode must be 'tiled' or 'direct'");
        } else if (key == "autotune") {
            autotune = parse_bool(value);
        } else if (key == "tile_order") {
            tile_order = parse_bool(value);
        } else if (key == "chunk_order") {
            chunk_order = parse_bool(value);
        } else if (key == "bin_accumulation") {
            if (value == "auto")
                bin_mode = 0;
            else if (value == "fixed")
                bin_mode = 1;
            else if (value == "float")
                bin_mode = 2;
            else
                fail(BCSIM_ERR_INVALID_ARGUMENT, "bin_accumulation must be auto, fixed or float");
        } else if (key == "fft_radix") {
            const int v = parse_int(value);
            if (v != 0 && v != 4 && v != 8) fail(BCSIM_ERR_INVALID_ARGUMENT, "fft_radix must be 4, 8 or 0 (automatic)");
            radix = v;
        } else if (key == "tile_windows" || key == "tile_lines" || key == "tile_block" || key == "tile_slices") {
            const int v = parse_int(value);
            if (v < 0) fail(BCSIM_ERR_INVALID_ARGUMENT, key + " must be >= 0 (0 = automatic)");
            if (key == "tile_lines" && v > 32) fail(BCSIM_ERR_INVALID_ARGUMENT, "tile_lines must be <= 32");
            if (key == "tile_block" && (v % 64 != 0 || v > 1024))
                fail(BCSIM_ERR_INVALID_ARGUMENT, "tile_block must be a multiple of 64, at most 1024");
            if (key == "tile_windows") tune.windows = v;
            if (key == "tile_lines") tune.lines = v;
            if (key == "tile_block") tune.block = v;
            if (key == "tile_slices") tune.slices = v;
        } else {
            fail(BCSIM_ERR_INVALID_ARGUMENT, "illegal parameter name: '" + key + "'");
        }
    }

    std::string get_parameter(const std::string& key) {
        if (key == "num_cuda_devices") {
            int count = 0;
            hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
            return std::to_string(count);
        }
        if (key == "cur_device_name") {
            hipDeviceProp_t prop;
            hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
            return std::string(prop.name) + " (" + prop.gcnArchName + ")";
        }
        if (key == "verbose") return std::to_string(verbose);
        if (key == "sound_speed") return std::to_string(sound_speed);
        if (key == "noise_amplitude") return std::to_string(noise_amplitude);
        if (key == "noise_seed") return std::to_string(noise_seed);
        if (key == "use_arc_projection") return arc ? "on" : "off";
        if (key == "radial_decimation") return std::to_string(decimation);
        if (key == "phase_delay") return phase_delay ? "on" : "off";
        if (key == "gpu_device") return std::to_string(device);
        if (key == "cuda_streams") return std::to_string(num_streams);
        if (key == "threads_per_block") return std::to_string(threads_per_block);
        if (key == "store_kernel_details") return store_details ? "on" : "off";
        if (key == "projection_counters") return count_pairs ? "on" : "off";
        if (key == "projection_mode") return direct_mode ? "direct" : "tiled";
        if (key == "weight_cutoff") {
            char b[32];
            std::snprintf(b, sizeof b, "%.9g", (double)weight_cutoff);
            return b;
        }
        if (key == "autotune") return autotune ? "on" : "off";
        if (key == "tile_order") return tile_order ? "on" : "off";
        if (key == "chunk_order") return chunk_order ? "on" : "off";
        if (key == "bin_accumulation") return bin_mode == 1 ? "fixed" : bin_mode == 2 ? "float" : "auto";
        if (key == "fft_radix") return std::to_string(radix);
        if (key == "launch_plan") return last_plan;
        if (key == "autotune_report") return TuneCache::get().report();
        if (key == "tile_windows") return std::to_string(tune.windows);
        if (key == "tile_lines") return std::to_string(tune.lines);
        if (key == "tile_block") return std::to_string(tune.block);
        if (key == "tile_slices") return std::to_string(tune.slices);
        fail(BCSIM_ERR_INVALID_ARGUMENT, "Illegal key: " + key);
    }

    // ---------------------------------------------------------------- configuration
    void set_excitation(const float* samples, int n, int center_index, float fs_, float fd) {
        ++version;
        if (!samples || n <= 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "excitation must have at least one sample");
        if (center_index < 0 || center_index >= n)
            fail(BCSIM_ERR_INVALID_ARGUMENT, "excitation center_index must lie inside the excitation");
        if (!(fs_ > 0)) fail(BCSIM_ERR_INVALID_ARGUMENT, "sampling frequency must be positive");
        can_change_device = false;
        exc.assign(samples, samples + n);
        center = center_index;
        fs = fs_;
        f_demod = fd;
        have_exc = true;
        dirty_scan = dirty_conv = dirty_basis = true;
        log(1, "Number of excitation samples: " + std::to_string(n));
    }

    void set_scan_sequence(int L, const float* o, const float* d, const float* lat, const float* t,
                           float length, bool equal) {
        ++version;
        if (!have_exc) fail(BCSIM_ERR_INVALID_STATE, "Excitation must be configured before scan sequence");
        if (L < 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "negative number of scan lines");
        if (L > 0 && (!o || !d || !lat || !t)) fail(BCSIM_ERR_INVALID_ARGUMENT, "null scan sequence array");
        std::vector<LineGeom> g(L);
        for (int i = 0; i < L; i++) {
            LineGeom& q = g[i];
            const float* oi = o + 3 * i;
            const float* di = d + 3 * i;
            const float* li = lat + 3 * i;
            q = LineGeom{oi[0], oi[1], oi[2], t[i], di[0], di[1], di[2], 0.f, li[0], li[1], li[2], 0.f,
                         0.f, 0.f, 0.f, 0.f};
            // elevational = lateral x direction (ScanSequence.cpp:40, vector3.hpp:96-101)
            q.ex = li[1] * di[2] - di[1] * li[2];
            q.ey = di[0] * li[2] - li[0] * di[2];
            q.ez = li[0] * di[1] - di[0] * li[1];
            const float e[3] = {q.ex, q.ey, q.ez};
            const float th = 1e-4f;  // ScanSequence.cpp:52-70
            bool ok = std::fabs(dot3(di, li)) < th && std::fabs(dot3(li, e)) < th && std::fabs(dot3(e, di)) < th;
            ok = ok && std::fabs((double)std::sqrt(dot3(di, di)) - 1.0) < th &&
                 std::fabs((double)std::sqrt(dot3(li, li)) - 1.0) < th &&
                 std::fabs((double)std::sqrt(dot3(e, e)) - 1.0) < th;
            if (!ok) fail(BCSIM_ERR_INVALID_ARGUMENT, "Invalid Scanline: Not orthogonal unit vectors (line " +
                                                           std::to_string(i) + ")");
        }
        lines.swap(g);
        line_length = length;
        all_equal = equal;
        have_scan = true;
        can_change_device = false;
        // the convolution tables follow the frame shape; prepare() rebuilds them
        // only if the new scan changes it
        dirty_scan = dirty_basis = true;
    }

    // Become a simulator of the same phantom as `src`: copy its configuration
    // (parameters, excitation, profile, scan) and share its resident datasets and
    // LUT; scan-dependent buffers, bins and tables are this simulator's own.
    void configure_view_of(const bcsim_sim& src) {
        ++version;
        verbose = src.verbose;
        sound_speed = src.sound_speed;
        noise_amplitude = src.noise_amplitude;
        noise_seed = src.noise_seed;
        arc = src.arc;
        decimation = src.decimation;
        phase_delay = src.phase_delay;
        device = src.device;
        direct_mode = src.direct_mode;
        tune = src.tune;
        autotune = src.autotune;
        radix = src.radix;
        weight_cutoff = src.weight_cutoff;
        bin_mode = src.bin_mode;
        have_exc = src.have_exc;
        exc = src.exc;
        center = src.center;
        fs = src.fs;
        f_demod = src.f_demod;
        have_scan = src.have_scan;
        lines = src.lines;
        line_length = src.line_length;
        all_equal = src.all_equal;
        fixed = src.fixed;
        splines = src.splines;
        spline_scan.clear();
        for (size_t i = 0; i < splines.size(); i++) spline_scan.push_back(std::make_unique<SplineScan>());
        profile = src.profile;
        sigma_l = src.sigma_l;
        sigma_e = src.sigma_e;
        lut = src.lut;
        lut_samples = src.lut_samples;
        lut_peak = src.lut_peak;
        lut_boxes = src.lut_boxes;
        lut_boxes_cut = src.lut_boxes_cut;
        lut_nr = src.lut_nr, lut_nl = src.lut_nl, lut_ne = src.lut_ne;
        lut_r0 = src.lut_r0, lut_r1 = src.lut_r1, lut_l0 = src.lut_l0;
        lut_l1 = src.lut_l1, lut_e0 = src.lut_e0, lut_e1 = src.lut_e1;
        can_change_device = false;
        dirty_scan = dirty_conv = dirty_basis = true;
    }

    // Give every line the time t (a frame of a dynamic phantom at one instant):
    // only the spline evaluation time changes, no table is rebuilt.
    void set_frame_time(float t) {
        ++version;
        for (auto& q : lines) q.t = t;
        all_equal = true;
        dirty_basis = true;
    }

    // Device bytes this simulator holds for its frames (not the shared datasets).
    size_t frame_bytes() const {
        size_t b = d_lines.bytes + d_H.bytes + d_demod.bytes + d_tw.bytes + d_bins.bytes + d_iq.bytes;
        for (auto& sc : spline_scan) b += sc->basis.bytes + sc->rendered.bytes;
        return b;
    }

    // The same before any allocation, from the frame shape (update_shape() done).
    size_t frame_bytes_estimate() const {
        const size_t L = lines.size();
        size_t b = L * sizeof(LineGeom) + (size_t)fft_len * sizeof(float2) * 2 + (size_t)n_out * sizeof(float2) +
                   L * (size_t)num_rf * (phase_delay ? 8 : 4);
        for (auto& sp : splines) b += sp->n * sizeof(float4) + (all_equal ? 0 : L * sizeof(SplineLineBasis));
        return b;
    }

    // Device bytes of the datasets and LUT (shared by views).
    size_t resident_bytes() const {
        size_t b = lut ? lut->bytes : 0;
        for (auto& f : fixed) b += f->pts.bytes + f->boxes.bytes;
        for (auto& sp : splines)
            b += sp->cx.bytes + sp->cy.bytes + sp->cz.bytes + sp->amp.bytes + sp->boxes.bytes + sp->iboxes.bytes;
        return b;
    }

    void clear_fixed() {
        quiesce();
        ++version;
        fixed.clear();
    }

    void add_fixed(const float* xyza, size_t n) {
        ++version;
        if (n > 0 && !xyza) fail(BCSIM_ERR_INVALID_ARGUMENT, "null scatterer array");
        if (n > (size_t)0x7fffffff) fail(BCSIM_ERR_INVALID_ARGUMENT, "too many scatterers in one dataset");
        use_device();
        quiesce();
        auto ds = std::make_shared<FixedDataset>();
        ds->n = n;
        ds->nchunks = (int)((n + kChunk - 1) / kChunk);
        amplitude_stats(xyza + 3, n, 4, ds->peak_amp, ds->typ_amp);
        const std::vector<uint32_t> perm = morton_order(xyza, n, 4);
        std::vector<float> sorted(n * 4);
        for (size_t i = 0; i < n; i++) std::memcpy(&sorted[4 * i], xyza + 4 * (size_t)perm[i], 4 * sizeof(float));
        std::vector<ChunkBox> boxes(ds->nchunks);
        for (int c = 0; c < ds->nchunks; c++) {
            BoxAcc acc;
            for (size_t i = (size_t)c * kChunk; i < std::min(n, (size_t)(c + 1) * kChunk); i++) acc.add(&sorted[4 * i]);
            boxes[c] = acc.box();
        }
        ds->visit.set_boxes(boxes);
        ds->pts.ensure(n * sizeof(float4));
        ds->boxes.ensure(boxes.size() * sizeof(ChunkBox));
        if (n) {
            hip_check(hipMemcpy(ds->pts.p, sorted.data(), n * sizeof(float4), hipMemcpyHostToDevice), "hipMemcpy");
            hip_check(hipMemcpy(ds->boxes.p, boxes.data(), boxes.size() * sizeof(ChunkBox), hipMemcpyHostToDevice),
                      "hipMemcpy");
        }
        fixed.push_back(std::move(ds));
        if (verbose) log(1, "Number of fixed scatterers: " + std::to_string(total_fixed()));
    }

    void clear_spline() {
        quiesce();
        ++version;
        splines.clear();
        spline_scan.clear();
    }

    void add_spline(int degree, const float* knots, int n_knots, const float* cps, const float* amps, size_t n,
                    int num_cp) {
        ++version;
        if (degree < 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "spline degree must be non-negative");
        if (degree > kMaxSplineDegree) fail(BCSIM_ERR_INVALID_ARGUMENT, "maximum spline degree supported is 4");
        if (num_cp <= degree) fail(BCSIM_ERR_INVALID_ARGUMENT, "too few spline control points for given degree");
        if (n_knots != num_cp + degree + 1)
            fail(BCSIM_ERR_INVALID_ARGUMENT, "mismatch in number of nodes, degree and control points");
        if (!knots || (n > 0 && (!cps || !amps))) fail(BCSIM_ERR_INVALID_ARGUMENT, "null spline array");
        if (n > (size_t)0x7fffffff) fail(BCSIM_ERR_INVALID_ARGUMENT, "too many scatterers in one dataset");
        use_device();
        quiesce();
        auto ds = std::make_shared<SplineDataset>();
        ds->degree = degree;
        ds->num_cp = num_cp;
        ds->n = n;
        ds->knots.assign(knots, knots + n_knots);
        ds->nchunks = (int)((n + kChunk - 1) / kChunk);
        amplitude_stats(amps, n, 1, ds->peak_amp, ds->typ_amp);
        // order by the centroid of each spline's control points
        std::vector<float> centroid(n * 3);
        for (size_t s = 0; s < n; s++) {
            double acc[3] = {0, 0, 0};
            for (int i = 0; i < num_cp; i++)
                for (int a = 0; a < 3; a++) acc[a] += cps[(s * num_cp + i) * 3 + a];
            for (int a = 0; a < 3; a++) centroid[3 * s + a] = (float)(acc[a] / num_cp);
        }
        const std::vector<uint32_t> perm = morton_order(centroid.data(), n, 3);
        // [n][num_cp][3] -> three control-point-major planes [num_cp][n], sorted
        std::vector<float> px((size_t)num_cp * n), py((size_t)num_cp * n), pz((size_t)num_cp * n), am(n);
        std::vector<ChunkBox> boxes(ds->nchunks);
        // per chunk and knot interval iv: the box of control points iv .. iv + degree,
        // the ones active on that interval
        const int n_iv = num_cp - degree;
        std::vector<ChunkBox> iboxes((size_t)ds->nchunks * n_iv);
        for (int c = 0; c < ds->nchunks; c++) {
            BoxAcc acc;
            std::vector<BoxAcc> iacc(n_iv);
            for (size_t s = (size_t)c * kChunk; s < std::min(n, (size_t)(c + 1) * kChunk); s++) {
                const size_t src = perm[s];
                am[s] = amps[src];
                for (int i = 0; i < num_cp; i++) {
                    const float* cp = cps + (src * num_cp + i) * 3;
                    px[(size_t)i * n + s] = cp[0];
                    py[(size_t)i * n + s] = cp[1];
                    pz[(size_t)i * n + s] = cp[2];
                    acc.add(cp);
                    for (int iv = std::max(0, i - degree); iv <= std::min(i, n_iv - 1); iv++) iacc[iv].add(cp);
                }
            }
            boxes[c] = acc.box();
            for (int iv = 0; iv < n_iv; iv++) iboxes[(size_t)c * n_iv + iv] = iacc[iv].box();
        }
        {
            double whole = 0.0, part = 0.0;
            for (int c = 0; c < ds->nchunks; c++) {
                const ChunkBox& b = boxes[c];
                if (!(b.hx >= 0.0f)) continue;
                whole += (double)b.hx + b.hy + b.hz;
                for (int iv = 0; iv < n_iv; iv++) {
                    const ChunkBox& q = iboxes[(size_t)c * n_iv + iv];
                    if (q.hx >= 0.0f) part += ((double)q.hx + q.hy + q.hz) / n_iv;
                }
            }
            ds->use_iboxes = part < 0.95 * whole;
        }
        const size_t b = px.size() * sizeof(float);
        ds->cx.ensure(b);
        ds->cy.ensure(b);
        ds->cz.ensure(b);
        ds->amp.ensure(n * sizeof(float));
        ds->visit.set_boxes(boxes);
        ds->boxes.ensure(boxes.size() * sizeof(ChunkBox));
        ds->iboxes.ensure(iboxes.size() * sizeof(ChunkBox));
        if (n) {
            hip_check(hipMemcpy(ds->iboxes.p, iboxes.data(), iboxes.size() * sizeof(ChunkBox), hipMemcpyHostToDevice),
                      "hipMemcpy");
            hip_check(hipMemcpy(ds->cx.p, px.data(), b, hipMemcpyHostToDevice), "hipMemcpy");
            hip_check(hipMemcpy(ds->cy.p, py.data(), b, hipMemcpyHostToDevice), "hipMemcpy");
            hip_check(hipMemcpy(ds->cz.p, pz.data(), b, hipMemcpyHostToDevice), "hipMemcpy");
            hip_check(hipMemcpy(ds->amp.p, am.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
            hip_check(hipMemcpy(ds->boxes.p, boxes.data(), boxes.size() * sizeof(ChunkBox), hipMemcpyHostToDevice),
                      "hipMemcpy");
        }
        splines.push_back(std::move(ds));
        spline_scan.push_back(std::make_unique<SplineScan>());
        dirty_basis = true;
        if (verbose) log(1, "Number of spline scatterers: " + std::to_string(total_spline()));
    }

    void set_gaussian(float sl, float se) {
        ++version;
        profile = kProfileGaussian;
        sigma_l = sl;
        sigma_e = se;
    }

    void set_lut(float r0, float r1, float l0, float l1, float e0, float e1, const float* samples, int nr,
                 int nl, int ne) {
        ++version;
        // LUTBeamProfile constructor (BeamProfile.cpp:58-72)
        if (nr <= 1) fail(BCSIM_ERR_INVALID_ARGUMENT, "Too few radial samples");
        if (nl <= 1) fail(BCSIM_ERR_INVALID_ARGUMENT, "Too few lateral samples");
        if (ne <= 1) fail(BCSIM_ERR_INVALID_ARGUMENT, "Too few elevational samples");
        if (!samples) fail(BCSIM_ERR_INVALID_ARGUMENT, "null LUT samples");
        use_device();
        quiesce();
        // quads of the (l, e) neighbourhood of slices r and r + 1, side by side (see
        // lut_weight in proj_math.h); the last slice, never a lower neighbour, pairs
        // with itself
        const size_t cells = (size_t)nr * nl * ne;
        std::vector<float4> quad(2 * cells);
        auto at = [&](int r, int l, int e) -> float {
            return (l < nl && e < ne) ? samples[((size_t)r * nl + l) * ne + e] : 0.0f;
        };
        auto q = [&](int r, int l, int e) {
            return make_float4(at(r, l, e), at(r, l, e + 1), at(r, l + 1, e), at(r, l + 1, e + 1));
        };
        for (int r = 0; r < nr; r++)
            for (int l = 0; l < nl; l++)
                for (int e = 0; e < ne; e++) {
                    const size_t cell = ((size_t)r * ne + e) * nl + l;
                    quad[2 * cell] = q(r, l, e);
                    quad[2 * cell + 1] = q(r + 1 < nr ? r + 1 : r, l, e);
                }
        auto table = std::make_shared<DevBuf>();
        table->ensure(quad.size() * sizeof(float4));
        hip_check(hipMemcpy(table->p, quad.data(), quad.size() * sizeof(float4), hipMemcpyHostToDevice),
                  "hipMemcpy");
        lut = std::move(table);
        lut_samples = std::make_shared<const std::vector<float>>(samples, samples + cells);
        lut_peak = 0.0f;
        for (size_t i = 0; i < cells; i++) lut_peak = std::max(lut_peak, std::fabs(samples[i]));
        lut_boxes.reset();
        lut_boxes_cut = -1.0f;
        lut_nr = nr;
        lut_nl = nl;
        lut_ne = ne;
        lut_r0 = r0; lut_r1 = r1;
        lut_l0 = l0; lut_l1 = l1;
        lut_e0 = e0; lut_e1 = e1;
        profile = kProfileLut;
    }

    size_t total_fixed() const {
        size_t s = 0;
        for (auto& f : fixed) s += f->n;
        return s;
    }
    size_t total_spline() const {
        size_t s = 0;
        for (auto& f : splines) s += f->n;
        return s;
    }

    // ---------------------------------------------------------------- derived tables
    int compute_num_rf() const {
        // compute_num_rf_samples<float> (common_utils.hpp:37-41)
        const float max_time = (float)(2.0 * (double)line_length / (double)sound_speed);
        const float prod = fs * max_time;
        const double v = std::floor((double)prod + 0.5);
        if (!(v >= 0.0 && v < 1e9)) fail(BCSIM_ERR_INVALID_ARGUMENT, "invalid line length");
        return (int)v;
    }

    void require_configured() {
        if (!have_exc) fail(BCSIM_ERR_INVALID_STATE, "Excitation not configured.");
        if (!have_scan) fail(BCSIM_ERR_INVALID_STATE, "Scan sequence not configured.");
    }

    void update_shape() {
        require_configured();
        const int nrf = compute_num_rf();
        if (nrf < 1) fail(BCSIM_ERR_INVALID_ARGUMENT, "line length gives no RF samples");
        size_t need = (size_t)nrf + exc.size() - 1;
        size_t N = 1;
        while (N < need) N <<= 1;  // next_power_of_two (fft.cpp:80-82)
        if (N > (size_t)kMaxFftLen) {
            log(3, "num_rf_samples required: " + std::to_string(nrf));
            fail(BCSIM_ERR_INVALID_ARGUMENT,
                 "Too many RF samples required: FFT length " + std::to_string(N) + " exceeds " +
                     std::to_string(kMaxFftLen));
        }
        num_rf = nrf;
        fft_len = (int)N;
        log2_len = 0;
        while ((1 << log2_len) < fft_len) log2_len++;
        n_out = (num_rf + decimation - 1) / decimation;
    }

    void upload_conv_tables() {
        const int N = fft_len;
        // H = mask * FFT_N(excitation) / N in double (BeamConvolver.cpp:77-91, the 1/N of
        // GpuAlgorithm.cpp:406), by direct summation over the excitation's few samples,
        // times exp(2 pi i k center / N): the spectrum of the output delayed by the
        // excitation's centre (CpuAlgorithm.cpp:353-358 reads sample center + D j).
        std::vector<float2> H(N);
        const double two_pi = 2.0 * M_PI;
        for (int k = 0; k < N; k++) {
            double mask = 0.0;
            if (N == 1) mask = 1.0;
            else if (k == 0) mask = 1.0;
            else if (k < N / 2) mask = 2.0;
            else if (k == N / 2) mask = 1.0;
            double re = 0, im = 0;
            if (mask != 0.0) {
                for (size_t n = 0; n < exc.size(); n++) {
                    const long long kn = ((long long)k * (long long)n) % N;
                    const double a = -two_pi * (double)kn / (double)N;
                    re += (double)exc[n] * std::cos(a);
                    im += (double)exc[n] * std::sin(a);
                }
            }
            const double sh = two_pi * (double)(((long long)k * center) % N) / (double)N;
            const double cs = std::cos(sh), sn = std::sin(sh);
            H[k] = make_float2((float)(mask * (re * cs - im * sn) / N), (float)(mask * (re * sn + im * cs) / N));
        }
        // demodulation: exp(-2 pi i nfd n), n = j*D, with nfd the reference's float quotient
        // (CpuAlgorithm.cpp:353-358); the cycle count nfd*n is exact in double.
        std::vector<float2> dem(n_out);
        const double nfd = (double)(f_demod / fs);
        for (int j = 0; j < n_out; j++) {
            double cyc = nfd * (double)((long long)j * decimation);
            cyc -= std::floor(cyc);
            const double a = -two_pi * cyc;
            dem[j] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        // two-level twiddles exp(-2 pi i k / N): coarse[k >> b] * fine[k & (2^b - 1)]
        const int b = log2_len / 2;
        const int nc = N >> b, nf = 1 << b;
        std::vector<float2> tw(nc + nf);
        for (int i = 0; i < nc; i++) {
            const double a = -two_pi * (double)((long long)i << b) / (double)N;
            tw[i] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        for (int i = 0; i < nf; i++) {
            const double a = -two_pi * (double)i / (double)N;
            tw[nc + i] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        d_H.ensure(sizeof(float2) * N);
        d_demod.ensure(sizeof(float2) * std::max(n_out, 1));
        d_tw.ensure(sizeof(float2) * tw.size());
        hip_check(hipMemcpy(d_H.p, H.data(), sizeof(float2) * N, hipMemcpyHostToDevice), "hipMemcpy");
        if (n_out) hip_check(hipMemcpy(d_demod.p, dem.data(), sizeof(float2) * n_out, hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(d_tw.p, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice), "hipMemcpy");
        tables_fft = fft_len;
        tables_out = n_out;
    }

    // Per spline dataset and line: knot interval and the p+1 float basis values
    // (GpuAlgorithm.cpp:657-692 does this per line with a constant-memory upload).
    void upload_spline_basis() {
        const int L = (int)lines.size();
        for (size_t q = 0; q < splines.size(); q++) {
            const SplineDataset* sp = splines[q].get();
            std::vector<SplineLineBasis> tab(std::max(L, 1));
            // all timestamps equal: the frame is rendered at line 0's time and the
            // table is not read, so only that time is checked
            const int checked = all_equal ? std::min(L, 1) : L;
            for (int k = 0; k < checked; k++) {
                const float t = lines[k].t;
                const int mu = knot_interval(sp->knots, t);
                const int lo = mu - sp->degree, hi = mu;
                if (lo < 0 || hi >= sp->num_cp)
                    fail(BCSIM_ERR_INVALID_ARGUMENT, "timestamp outside the span of the spline basis");
                SplineLineBasis& B = tab[k];
                std::memset(&B, 0, sizeof(B));
                B.lo = lo;
                B.count = hi - lo + 1;
                for (int i = 0; i < B.count; i++) B.b[i] = bspline_basis(lo + i, sp->degree, t, sp->knots);
            }
            if (all_equal) continue;
            DevBuf& basis = spline_scan[q]->basis;
            basis.ensure(sizeof(SplineLineBasis) * tab.size());
            hip_check(hipMemcpy(basis.p, tab.data(), sizeof(SplineLineBasis) * tab.size(),
                                hipMemcpyHostToDevice), "hipMemcpy");
        }
    }

    void prepare() {
        require_configured();
        if (lines.empty()) fail(BCSIM_ERR_INVALID_STATE, "No scanlines in scansequence");
        if (profile < 0) fail(BCSIM_ERR_INVALID_STATE, "No beam profile is configured");
        use_device();
        if (dirty_scan || dirty_conv || dirty_basis) quiesce();
        if (dirty_scan || dirty_conv) {
            // update_shape() may already have run (bcsim_get_output_shape): compare the
            // shape with the one the tables were built for, not with the previous shape
            update_shape();
            if (dirty_conv || fft_len != tables_fft || n_out != tables_out) upload_conv_tables();
            d_lines.ensure(sizeof(LineGeom) * lines.size());
            hip_check(hipMemcpy(d_lines.p, lines.data(), sizeof(LineGeom) * lines.size(), hipMemcpyHostToDevice),
                      "hipMemcpy");
            dirty_scan = dirty_conv = false;
        }
        if (dirty_basis) {
            upload_spline_basis();
            dirty_basis = false;
        }
        if (profile == kProfileLut && weight_cutoff > 0.0f && (!lut_boxes || lut_boxes_cut != weight_cutoff)) {
            quiesce();
            build_slice_boxes();
        }
        const size_t L = lines.size();
        const size_t bin_bytes = L * (size_t)num_rf * (phase_delay ? 8 : 4);
        if (bin_bytes > d_bins.bytes) quiesce();
        d_bins.ensure(bin_bytes);
    }

    // Sparse table of the LUT's per-slice record boxes for the weight cutoff (see
    // ProfileParams::slice_box): level 0 holds, for each radial record index r, the
    // ranges of (l, e) whose record -- samples (r..r+1, l..l+1, e..e+1), the quad
    // pair lut_weight reads -- has a sample of magnitude >= cutoff * peak; level k
    // the union over records [r, r + 2^k).
    void build_slice_boxes() {
        const int nr = lut_nr, nl = lut_nl, ne = lut_ne;
        const std::vector<float>& v = *lut_samples;
        const float thr = weight_cutoff * lut_peak;
        const int kEmpty = 1 << 29;
        std::vector<uint8_t> big((size_t)nr * nl * ne);
        for (size_t i = 0; i < big.size(); i++) big[i] = std::fabs(v[i]) >= thr;
        auto at = [&](int r, int l, int e) -> bool {
            return l < nl && e < ne && big[((size_t)r * nl + l) * ne + e];
        };
        int levels = 1;
        while ((1 << levels) <= nr) levels++;
        std::vector<int4> st((size_t)levels * nr);
        for (int r = 0; r < nr; r++) {
            const int r1 = r + 1 < nr ? r + 1 : r;
            int4 b = make_int4(kEmpty, -1, kEmpty, -1);
            for (int l = 0; l < nl; l++)
                for (int e = 0; e < ne; e++) {
                    bool any = false;
                    for (int q : {r, r1})
                        any = any || at(q, l, e) || at(q, l, e + 1) || at(q, l + 1, e) || at(q, l + 1, e + 1);
                    if (!any) continue;
                    b.x = std::min(b.x, l), b.y = std::max(b.y, l);
                    b.z = std::min(b.z, e), b.w = std::max(b.w, e);
                }
            st[r] = b;
        }
        for (int k = 1; k < levels; k++)
            for (int r = 0; r < nr; r++) {
                const int4 a = st[(size_t)(k - 1) * nr + r];
                const int r2 = r + (1 << (k - 1));
                const int4 c = r2 < nr ? st[(size_t)(k - 1) * nr + r2] : a;
                st[(size_t)k * nr + r] = make_int4(std::min(a.x, c.x), std::max(a.y, c.y), std::min(a.z, c.z),
                                                   std::max(a.w, c.w));
            }
        auto buf = std::make_shared<DevBuf>();
        buf->ensure(st.size() * sizeof(int4));
        hip_check(hipMemcpy(buf->p, st.data(), st.size() * sizeof(int4), hipMemcpyHostToDevice), "hipMemcpy");
        lut_boxes = std::move(buf);
        lut_boxes_cut = weight_cutoff;
    }

    ProjCommon proj_common(bool spline_rule) const {
        ProjCommon c{};
        c.lines = d_lines.as<LineGeom>();
        c.num_lines = (int)lines.size();
        c.num_rf = num_rf;
        if (spline_rule) {
            // r*2.0/(c*float(1/fs)) (CpuAlgorithm.cpp:161-162)
            const float dt = (float)(1.0 / (double)fs);
            c.bin_scale = 2.0;
            c.bin_denom = (double)(sound_speed * dt);
        } else {
            // r*2.0*fs/c (CpuAlgorithm.cpp:75)
            c.bin_scale = 2.0 * (double)fs;
            c.bin_denom = (double)sound_speed;
        }
        c.bin_inv_denom = 1.0 / c.bin_denom;
        c.phase_per_sample = (float)(2.0 * M_PI * (double)f_demod / (double)fs);
        ProfileParams& P = c.prof;
        if (profile == kProfileGaussian) {
            P.inv_tsl = (float)(1.0 / (2.0 * (double)sigma_l * (double)sigma_l));
            P.inv_tse = (float)(1.0 / (2.0 * (double)sigma_e * (double)sigma_e));
            P.g_cut = weight_cutoff > 0.0f ? (float)(-std::log((double)weight_cutoff)) : INFINITY;
        } else {
            P.g_cut = INFINITY;
            const bool cut = weight_cutoff > 0.0f && lut_boxes && lut_boxes_cut == weight_cutoff;
            P.slice_box = cut ? lut_boxes->as<int4>() : nullptr;
            P.w_cut = cut ? weight_cutoff * lut_peak : 0.0f;
            P.quad = lut->as<float4>();
            P.nr = lut_nr;
            P.nl = lut_nl;
            P.ne = lut_ne;
            P.r_min = lut_r0;
            P.l_min = lut_l0;
            P.e_min = lut_e0;
            // grid spacing as LUTBeamProfile stores it: a float quotient (BeamProfile.cpp:68-70)
            P.inv_dr = (float)(1.0 / (double)((lut_r1 - lut_r0) / (float)(lut_nr - 1)));
            P.inv_dl = (float)(1.0 / (double)((lut_l1 - lut_l0) / (float)(lut_nl - 1)));
            P.inv_de = (float)(1.0 / (double)((lut_e1 - lut_e0) / (float)(lut_ne - 1)));
        }
        return c;
    }

    // Launch shape of the projection: a tile of lines whose bins fit in LDS, and
    // enough scatterer slices to give the 256 CUs a few workgroups each.
    ProjLaunch proj_launch(size_t n) const {
        ProjLaunch cfg{};
        const size_t per_line = (size_t)num_rf * (phase_delay ? 8 : 4);
        int tile = (int)std::min<size_t>(8, (size_t)(kLdsBytes - 1024) / per_line);
        tile = std::max(1, std::min(tile, (int)lines.size()));
        cfg.tile_lines = tile;
        const int tiles = ((int)lines.size() + tile - 1) / tile;
        const int target = 2048;
        int slices = std::max(1, (target + tiles - 1) / tiles);
        const int max_slices = std::max<int>(1, (int)((n + 2047) / 2048));
        cfg.slices = std::min(slices, max_slices);
        cfg.block = 256;
        return cfg;
    }

    // ---------------------------------------------------------------- frame
    struct Timed {
        bcsim_sim* s;
        const char* key;
        const char* alias;  // a second key that records the same duration, or null
        hipEvent_t a = nullptr, b = nullptr;
        Timed(bcsim_sim* s_, const char* k, const char* al = nullptr) : s(s_), key(k), alias(al) {
            if (!s->store_details) return;
            hip_check(hipEventCreate(&a), "hipEventCreate");
            hip_check(hipEventCreate(&b), "hipEventCreate");
            hip_check(hipEventRecord(a, s->stream), "hipEventRecord");
        }
        void stop() {
            if (!a) return;
            hip_check(hipEventRecord(b, s->stream), "hipEventRecord");
            s->pending.push_back({key, alias ? alias : "", a, b});
            a = b = nullptr;
        }
        ~Timed() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    };
    struct Pending {
        std::string key, alias;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;

    size_t stats_launches = 0;  // projection launches of the last frame that filled d_stats
    static constexpr size_t kStatsPerLaunch = 3;  // TileLaunch::stats

    void collect_timings() {
        if (stats_launches) {
            // the launches' own counts (TileLaunch::stats): the weight evaluations and
            // LDS additions of each projection launch
            std::vector<unsigned long long> st(kStatsPerLaunch * stats_launches);
            hip_check(hipMemcpyAsync(st.data(), d_stats.p, st.size() * sizeof(unsigned long long),
                                     hipMemcpyDeviceToHost, stream),
                      "hipMemcpyAsync");
            hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            for (size_t i = 0; i < stats_launches; i++) {
                debug["projection_weight_evals"].push_back((double)st[kStatsPerLaunch * i]);
                debug["projection_lds_adds"].push_back((double)st[kStatsPerLaunch * i + 1]);
                debug["projection_chunk_line_pairs"].push_back((double)st[kStatsPerLaunch * i + 2]);
            }
            stats_launches = 0;
        }
        for (auto& p : pending) {
            float ms = 0.f;
            if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                debug[p.key].push_back((double)ms);
                if (!p.alias.empty()) debug[p.alias].push_back((double)ms);
            }
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        pending.clear();
    }

    // Beam support for the chunk cull (see CullParams): the radial window that
    // gives a bin in [0, num_rf), and the profile's non-zero region.
    CullParams cull_params(const ProjCommon& c) const {
        CullParams cp{};
        const double per_sample = c.bin_denom / c.bin_scale;  // metres per bin
        double r_lo = -0.5 * per_sample, r_hi = ((double)num_rf - 0.5) * per_sample;
        cp.l_lo = cp.e_lo = -INFINITY;
        cp.l_hi = cp.e_hi = INFINITY;
        cp.g_a = cp.g_b = 0.0f;
        cp.g_max = INFINITY;
        if (profile == kProfileLut) {
            const double dr = (double)((lut_r1 - lut_r0) / (float)(lut_nr - 1));
            const double dl = (double)((lut_l1 - lut_l0) / (float)(lut_nl - 1));
            const double de = (double)((lut_e1 - lut_e0) / (float)(lut_ne - 1));
            r_lo = std::max(r_lo, (double)lut_r0 - std::fabs(dr));
            r_hi = std::min(r_hi, (double)lut_r0 + (lut_nr - 1) * std::fabs(dr));
            cp.l_lo = (float)((double)lut_l0 - std::fabs(dl));
            cp.l_hi = (float)((double)lut_l0 + (lut_nl - 1) * std::fabs(dl));
            cp.e_lo = (float)((double)lut_e0 - std::fabs(de));
            cp.e_hi = (float)((double)lut_e0 + (lut_ne - 1) * std::fabs(de));
            cp.lut_dl = (float)dl;
            cp.lut_de = (float)de;
            if (!(dr > 0 && dl > 0 && de > 0)) {  // degenerate table: no culling
                r_lo = -INFINITY;
                r_hi = INFINITY;
                cp.l_lo = cp.e_lo = -INFINITY;
                cp.l_hi = cp.e_hi = INFINITY;
                cp.lut_dl = cp.lut_de = 0.0f;  // and no narrowing by the weight cutoff's boxes
            }
        } else {
            // expf(-x) rounds to zero for x > 103.98 (below half the smallest denormal);
            // 110 leaves a 6 % margin on the exponent.  With a weight cutoff the pairs
            // beyond g_cut weigh zero; the cull keeps a relative 1e-4 above it, far
            // more than the rounding of a pair's exponent.
            cp.g_a = c.prof.inv_tsl;
            cp.g_b = c.prof.inv_tse;
            float gm = 110.0f;
            if (std::isfinite(c.prof.g_cut)) gm = std::min(gm, c.prof.g_cut * 1.0001f);
            cp.g_max = (std::isfinite(cp.g_a) && std::isfinite(cp.g_b)) ? gm : INFINITY;
        }
        cp.r_per_bin = (float)per_sample;
        cp.r_lo = (float)r_lo;
        cp.r_hi = (float)r_hi;
        return cp;
    }

    bool shared_origin() const {
        for (const auto& q : lines)
            if (std::memcmp(&q.ox, &lines[0].ox, 3 * sizeof(float)) != 0) return false;
        return true;
    }

    // Launch shape of the tile kernel for `tune` (0 fields: heuristic): one
    // radial window unless a line's bins overflow LDS, up to 16 lines per
    // workgroup (while at least ~1024 workgroups remain), and enough chunk
    // slices for about 256 workgroups (1024 for splines with a time per line,
    // whose hits cost more: the tuner's choices on the 2D sector and the spline
    // M-mode, 6 % and 13 % faster than 512 workgroups).
    TileLaunch tile_launch(int nchunks, const TileTune& tune, bool per_line_splines, bool fixed_bins) const {
        TileLaunch cfg{};
        const int L = (int)lines.size();
        const size_t bin_bytes = tile_bin_bytes(phase_delay, fixed_bins);
        const size_t avail = (size_t)kLdsBytes - 64 * 8 - 256;  // keep room for a 64-entry list
        // radial windows: one by default (measured fastest on the workloads of the tests
        // and bench); more if one line's bins (and its geometry, kTileLineLds) do not fit
        int K = tune.windows > 0 ? tune.windows : 1;
        K = std::max(1, std::min(K, num_rf));
        while ((size_t)((num_rf + K - 1) / K) * bin_bytes + kTileLineLds > avail) K++;
        const int W = (num_rf + K - 1) / K;
        K = (num_rf + W - 1) / W;
        const size_t per_line = (size_t)W * bin_bytes + kTileLineLds;
        const int tmax = (int)std::max<size_t>(1, std::min<size_t>(32, avail / per_line));
        int T;
        if (tune.lines > 0) {
            T = std::min(tune.lines, tmax);
        } else {
            // as many lines as fit, but keep at least ~1024 workgroups
            T = std::min(tmax, 16);
            const int want = std::max(1, (L * K) / 1024);
            T = std::max(1, std::min(T, want));
        }
        T = std::max(1, std::min(T, L));
        const int blocks = ((L + T - 1) / T) * K;
        const int target = per_line_splines ? 1024 : 256;
        int slices = tune.slices > 0 ? tune.slices : std::max(1, (target + blocks - 1) / blocks);
        slices = std::max(1, std::min(slices, std::max(1, nchunks / 16)));
        cfg.tile_lines = T;
        cfg.windows = K;
        cfg.window_bins = W;
        cfg.chunks_per_slice = std::max(1, (nchunks + slices - 1) / slices);
        cfg.slices = std::max(1, (nchunks + cfg.chunks_per_slice - 1) / cfg.chunks_per_slice);
        cfg.block = tune.block > 0 ? tune.block : 1024;
        cfg.atomic = cfg.slices > 1 ? 1 : 0;
        cfg.fixed = fixed_bins ? 1 : 0;
        // the rest of the LDS holds the chunk list.  When the LDS would hold it below
        // its cap and a tile has at most 4 lines, an entry is one word (chunk << 4 |
        // line mask) instead of two: twice the entries, so the list drains half as
        // often (1-GPU volume, 4 lines per tile: 896 -> 1792 entries, projection
        // 1.644 -> 1.546 ms; with phase delay 2.665 -> 2.537 ms)
        const size_t left = (size_t)kLdsBytes - 256 - (size_t)T * per_line;
        const bool packable = T <= 4 && nchunks < (1 << 28);
        cfg.list_packed = packable && (left / 8) / 64 * 64 < (size_t)kMaxListCap ? 1 : 0;
        cfg.list_cap = (int)std::min<size_t>(kMaxListCap, (left / (cfg.list_packed ? 4 : 8)) / 64 * 64);
        return cfg;
    }

    // ---------------------------------------------------------------- launch planning
    // One projection launch of the frame: a dataset and how it is projected.
    struct ProjTarget {
        const char* kind;  // "fixed", "rendered" (spline at one time) or "spline"
        ProjCommon c;
        FixedSet fs{};
        SplineSet ss{};
        const ChunkBox* boxes = nullptr;
        int nchunks = 0;
        size_t n = 0;
        const SplineDataset* sp = nullptr;  // spline kinds
        SplineScan* scan = nullptr;
        const int* chunk_order = nullptr;   // visiting order of the chunks (ChunkOrders), or null
        const float* chunk_dist = nullptr;  // their reach from the common origin (ChunkOrders), or null
    };

    // Largest |weight| of the beam profile: 1 for the Gaussian, the largest |sample| of a LUT.
    double profile_peak() const { return profile == kProfileLut ? (double)lut_peak : 1.0; }

    // Does this target's launch use fixed-point bins?  Not where its amplitudes do
    // not allow it (amp_scale 0); else the parameter if set, then `tuned` (a
    // TileTune::bins), then the default: fixed for the Gaussian profile, whose LDS
    // additions weigh on the kernel (fixed-point ones are ~10x cheaper), float for
    // a LUT, whose table gathers bound it and where the wider bins only halve the
    // lines a tile holds (measured on the 1M-point volume: 1.91 against 1.65 ms).
    bool target_fixed(const ProjTarget& t, int tuned = 0) const {
        if (!(t.c.amp_scale > 0.0f)) return false;
        const int m = bin_mode ? bin_mode : tuned ? tuned : (profile == kProfileLut ? 2 : 1);
        return m == 1;
    }

    // Key of this scan's chunk visiting order (see ChunkOrders).
    ChunkOrders::Key chunk_order_key() const {
        if (lines.empty()) return ChunkOrders::Key{};
        if (shared_origin()) return ChunkOrders::Key{lines[0].ox, lines[0].oy, lines[0].oz, 0.f, 0.f, 0.f, 1.f};
        const LineGeom& m = lines[lines.size() / 2];
        return ChunkOrders::Key{m.ox, m.oy, m.oz, m.rx, m.ry, m.rz, 0.f};
    }

    // Where an allocation-free frame keeps what a frame otherwise holds in buffers
    // of the simulator: the bins, then the rendered points of all-timestamps-equal
    // spline sets, carved from the fixed scratch array (256-byte aligned pieces).
    struct FrameMem {
        unsigned char* scratch = nullptr;
        size_t used = 0, cap = 0;
        float* bins = nullptr;
        void* carve(size_t bytes) {
            const size_t at = (used + 255) & ~(size_t)255;
            if (at + bytes > cap) fail(BCSIM_ERR_CAPACITY, "allocation-free frame: scratch exhausted");
            used = at + bytes;
            return scratch + at;
        }
    };

    // Bytes of scratch an allocation-free frame of the current configuration needs.
    size_t noalloc_bytes() const {
        size_t b = (lines.size() * (size_t)num_rf * (phase_delay ? 8 : 4) + 255) & ~(size_t)255;
        if (all_equal)
            for (auto& sp : splines) b += (sp->n * sizeof(float4) + 255) & ~(size_t)255;
        return b;
    }

    // The frame's projection launches.  fm (allocation-free frames): rendered
    // points go to its scratch and only chunk orders built earlier are used.
    std::vector<ProjTarget> proj_targets(FrameMem* fm = nullptr) {
        std::vector<ProjTarget> out;
        const ChunkOrders::Key okey = chunk_order_key();
        const bool create = fm == nullptr;
        for (auto& f : fixed) {
            if (f->n == 0) continue;
            ProjTarget t{"fixed", proj_common(false)};
            set_fixed_scale(t.c, f->peak_amp * profile_peak(), f->typ_amp * profile_peak(), f->n);
            t.fs = FixedSet{f->pts.as<float4>(), (int)f->n};
            t.boxes = f->boxes.as<ChunkBox>();
            t.nchunks = f->nchunks;
            t.n = f->n;
            if (chunk_order) {
                t.chunk_order = f->visit.get(okey, create);
                t.chunk_dist = ChunkOrders::distances(okey, t.chunk_order, t.nchunks);
            }
            out.push_back(t);
        }
        for (size_t q = 0; q < splines.size(); q++) {
            const SplineDataset* sp = splines[q].get();
            SplineScan* scan = spline_scan[q].get();
            if (sp->n == 0) continue;
            ProjTarget t{all_equal ? "rendered" : "spline", proj_common(true)};
            set_fixed_scale(t.c, sp->peak_amp * profile_peak(), sp->typ_amp * profile_peak(), sp->n);
            t.ss = SplineSet{sp->cx.as<float>(), sp->cy.as<float>(), sp->cz.as<float>(), sp->amp.as<float>(),
                             scan->basis.as<SplineLineBasis>(), (int)sp->n, sp->num_cp,
                             sp->use_iboxes ? sp->iboxes.as<ChunkBox>() : nullptr, sp->num_cp - sp->degree};
            if (all_equal) {
                float4* dst;
                if (fm) {
                    dst = static_cast<float4*>(fm->carve(sp->n * sizeof(float4)));
                } else {
                    scan->rendered.ensure(sp->n * sizeof(float4));
                    dst = scan->rendered.as<float4>();
                }
                t.fs = FixedSet{dst, (int)sp->n};
            }
            t.boxes = sp->boxes.as<ChunkBox>();
            t.nchunks = sp->nchunks;
            t.n = sp->n;
            t.sp = sp;
            t.scan = scan;
            if (chunk_order) {
                t.chunk_order = sp->visit.get(okey, create);
                t.chunk_dist = ChunkOrders::distances(okey, t.chunk_order, t.nchunks);
            }
            out.push_back(t);
        }
        return out;
    }

    // Render the all-timestamps-equal spline sets at line 0's time
    // (GpuAlgorithm.cpp:181-187, GpuScatterers.cpp:48-123).
    void enqueue_renders(const std::vector<ProjTarget>& targets) {
        if (!all_equal) return;
        for (const ProjTarget& t : targets) {
            Timed tm(this, "kernel_render_ms");
            render_target(t);
            tm.stop();
        }
    }

    void render_target(const ProjTarget& t) {
        if (std::strcmp(t.kind, "rendered") != 0) return;
        SplineLineBasis B0;
        host_basis_line0(*t.sp, B0);
        hip_check(launch_render_spline(t.ss, B0, const_cast<float4*>(t.fs.pts), stream), "render kernel");
    }

    // Blocks (tile, window) of one slice of a tile launch.
    int tile_blocks(const TileLaunch& cfg) const {
        return (((int)lines.size() + cfg.tile_lines - 1) / cfg.tile_lines) * cfg.windows;
    }

    // Fewer blocks than this start together on the 256 CUs: no order to choose.
    static constexpr int kMinOrderedBlocks = 512;

    // [work | order] for `blocks` blocks of projection target i (zeroed on the
    // stream when the buffer is new; afterwards the order kernel clears the work
    // counts every frame).
    int* order_buffer(size_t i, int blocks) {
        if (tile_order_bufs.size() <= i) tile_order_bufs.resize(i + 1);
        auto& b = tile_order_bufs[i];
        if (!b) b = std::make_unique<DevBuf>();
        const size_t need = 2 * (size_t)blocks * sizeof(int);
        if (b->bytes < need) {
            b->ensure(need);
            hip_check(launch_fill_zero(b->as<float>(), b->bytes / sizeof(float), stream), "fill kernel");
        }
        return b->as<int>();
    }

    void launch_target(const ProjTarget& t, const TileLaunch& cfg_in, float* bins) {
        const bool is_spline = std::strcmp(t.kind, "spline") == 0;
        TileLaunch cfg = cfg_in;
        cfg.chunks = t.chunk_order;
        cfg.chunk_dist = t.chunk_dist;
        hip_check(launch_project_tile(t.c, cull_params(t.c), is_spline ? nullptr : &t.fs, is_spline ? &t.ss : nullptr,
                                      t.boxes, t.nchunks, bins, arc, phase_delay, profile == kProfileLut,
                                      shared_origin() && !is_spline, cfg, stream),
                  is_spline ? "spline projection kernel" : "fixed projection kernel");
    }

    std::string shape_class(const char* kind, int nchunks, bool fixed_ok) const {
        char buf[160];
        std::snprintf(buf, sizeof buf, "tile/%s/%s/%s/bins-%s/%s/L2^%d/C2^%d/F%d", kind,
                      profile == kProfileLut ? "lut" : "gauss", phase_delay ? "complex" : "real",
                      !fixed_ok ? "float-only" : bin_mode == 1 ? "fixed" : bin_mode == 2 ? "float" : "auto",
                      arc ? "arc" : "noarc", ceil_log2((long long)lines.size()), ceil_log2(nchunks), fft_len);
        return buf;
    }

    // Tuned tile shape of this target's class: looked up, or found now by a
    // coordinate search from the heuristic shape (bin accumulation and radial
    // windows, then lines per tile, then total workgroups, then block size), each
    // candidate timed on this frame's data.
    // Writes the bins: call before they are cleared for the frame.
    TileTune tuned_tile(const ProjTarget& t) {
        const bool fixed_ok = t.c.amp_scale > 0.0f;
        const std::string key = shape_class(t.kind, t.nchunks, fixed_ok);
        TileTune best;
        if (TuneCache::get().find_tile(key, best)) return best;
        float* bins = d_bins.as<float>();
        render_target(t);  // trials project the points of this frame's time
        std::string trials;
        const bool spl = std::strcmp(t.kind, "spline") == 0;
        // each candidate runs as a frame would: for launches of many blocks, its
        // blocks heaviest first by the counts of its own previous launch, so two warm
        // launches settle the order before three are timed
        auto cost = [&](const TileTune& tt) {
            TileLaunch cfg = tile_launch(t.nchunks, tt, spl, target_fixed(t, tt.bins));
            const int blocks = tile_blocks(cfg);
            const bool ordered = tile_order && blocks >= kMinOrderedBlocks && blocks <= kMaxOrderedTiles;
            DevBuf order;
            if (ordered) {
                order.ensure(2 * (size_t)blocks * sizeof(int));
                hip_check(launch_fill_zero(order.as<float>(), 2 * (size_t)blocks, stream), "fill kernel");
                cfg.work = order.as<int>();
                cfg.order = order.as<int>() + blocks;
            }
            const double ms = time_launches(
                stream,
                [&] {
                    if (ordered)
                        hip_check(launch_order_tiles(cfg.work, const_cast<int*>(cfg.order), blocks, stream),
                                  "tile order kernel");
                    launch_target(t, cfg, bins);
                },
                3, 2);
            char b[64];
            std::snprintf(b, sizeof b, " [%d,%d,%d,%d,%s]=%.3f", cfg.tile_lines, cfg.block, cfg.slices, cfg.windows,
                          cfg.fixed ? "fixed" : "float", ms);
            trials += b;
            return ms;
        };
        // the bin accumulation (unless set or not possible) and the radial windows
        // first, each pair at its heuristic lines per tile: fixed-point bins take twice
        // the LDS, and windows give back lines per tile (measured on the 1M-point volume
        // with phase delay: fixed bins in 4 windows of 4 lines, 2.50 ms of projection
        // against 2.67 for float bins in whole lines of 2)
        best.bins = target_fixed(t) ? 1 : 2;
        double best_ms = cost(best);
        std::vector<int> modes{best.bins};
        if (fixed_ok && !bin_mode) modes.push_back(best.bins == 1 ? 2 : 1);
        for (int b : modes) {
            for (int K : {1, 2, 4}) {
                if (b == best.bins && K == 1 && best.windows == 0) continue;  // the heuristic shape, timed above
                TileTune c;
                c.bins = b;
                c.windows = K;
                const double ms = cost(c);
                if (ms < best_ms) best = c, best_ms = ms;
            }
        }
        const bool fx = target_fixed(t, best.bins);
        TileTune wide = best;  // the most lines a tile of the chosen bins and windows holds
        wide.lines = 32;
        const int tmax = tile_launch(t.nchunks, wide, spl, fx).tile_lines;
        for (int T = 1; T <= tmax; T *= 2) {
            TileTune c = best;
            c.lines = T;
            const double ms = cost(c);
            if (ms < best_ms) best = c, best_ms = ms;
        }
        const TileLaunch cur = tile_launch(t.nchunks, best, spl, fx);  // (best.bins stays as chosen)
        const int blocks = (((int)lines.size() + cur.tile_lines - 1) / cur.tile_lines) * cur.windows;
        for (int wgs : {256, 512, 1024, 2048, 4096}) {
            TileTune c = best;
            c.lines = cur.tile_lines;
            c.slices = std::max(1, (wgs + blocks - 1) / blocks);
            if (c.slices == cur.slices) continue;
            const double ms = cost(c);
            if (ms < best_ms) best = c, best_ms = ms;
        }
        for (int b : {256, 512}) {
            TileTune c = best;
            c.block = b;
            const double ms = cost(c);
            if (ms < best_ms) best = c, best_ms = ms;
        }
        char head[64];
        std::snprintf(head, sizeof head, "%.3f ms with ", best_ms);
        TuneCache::get().put_tile(key, best,
                                  head + describe(best) + "; trials [lines,block,slices,windows,bins]:" + trials);
        return best;
    }

    // Launch shape for a target: shape fields set by the user win, then the
    // tuned shape (when tuning is on), then the heuristic.
    TileLaunch plan_tile(const ProjTarget& t, bool tune_now) {
        TileTune eff = tune;
        if (tune_now) {
            const TileTune tt = tuned_tile(t);
            if (!eff.windows) eff.windows = tt.windows;
            if (!eff.lines) eff.lines = tt.lines;
            if (!eff.block) eff.block = tt.block;
            if (!eff.slices) eff.slices = tt.slices;
            eff.bins = tt.bins;
        }
        return tile_launch(t.nchunks, eff, std::strcmp(t.kind, "spline") == 0, target_fixed(t, eff.bins));
    }

    // FFT pass radix of this frame's convolution: set by the user, tuned per
    // (FFT length, bin type, line count class), or the default.
    int plan_radix(bool tune_now) {
        if (radix) return radix;
        if (!tune_now || lines.empty()) return kDefaultRadix;
        char key[96];
        std::snprintf(key, sizeof key, "conv/F%d/%s/L2^%d", fft_len, phase_delay ? "complex" : "real",
                      ceil_log2((long long)lines.size()));
        int r = 0;
        if (TuneCache::get().find_radix(key, r)) return r;
        DevBuf scratch;  // trial output, freed after the trials
        scratch.ensure(lines.size() * (size_t)n_out * sizeof(float2));
        const uint32_t counter = noise_counter;  // trials do not advance the noise stream
        ConvParams p = conv_params(scratch.as<float2>());
        noise_counter = counter;
        std::string trials;
        double best_ms = 0.0;
        for (int cand : {4, 8}) {
            p.radix = cand;
            const double ms =
                time_launches(stream, [&] { hip_check(launch_convolve_demod(p, stream), "convolve"); }, 3, 2);
            char b[32];
            std::snprintf(b, sizeof b, " r%d=%.3f", cand, ms);
            trials += b;
            if (r == 0 || ms < best_ms) r = cand, best_ms = ms;
        }
        TuneCache::get().put_radix(key, r, "radix " + std::to_string(r) + ";" + trials);
        return r;
    }

    // Debug data keys the reference records on every frame with
    // store_kernel_details (GpuAlgorithm.cpp:193-358) that have no kernel of their
    // own here: one stream per frame (stream number 0 for every line), and the
    // multiply, inverse transform and demodulation fused into the convolution
    // kernel, whose duration is kernel_forward_fft_ms (and
    // kernel_convolve_demodulate_ms).
    void record_fused_keys() {
        debug["stream_numbers"].assign(lines.size(), 0.0);
        for (const char* k : {"kernel_multiply_fft_ms", "kernel_inverse_fft_ms", "kernel_demodulate_ms"})
            debug[k].push_back(0.0);
    }

    // Enqueue one frame.  fm: an allocation-free frame (its bins and rendered points
    // in the fixed scratch array; heuristic or user-set launch shapes, no tuning
    // trials, no block-order buffers, no timing events).
    void enqueue_frame(float2* d_out, FrameMem* fm = nullptr) {
        if (!fm) prepare();
        if (store_details || count_pairs) debug.clear();
        if (store_details) record_fused_keys();
        const bool cplx = phase_delay;
        const size_t L = lines.size();
        const size_t bin_floats = L * (size_t)num_rf * (cplx ? 2 : 1);
        float* bins = fm ? fm->bins : d_bins.as<float>();
        const bool lut_profile = profile == kProfileLut;

        const std::vector<ProjTarget> targets = proj_targets(fm);
        enqueue_renders(targets);
        // plan every launch first: tuning trials write the bins
        std::vector<TileLaunch> plans;
        std::vector<ProjLaunch> direct_plans;
        last_plan.clear();
        for (const ProjTarget& t : targets) {
            if (direct_mode) {
                direct_plans.push_back(proj_launch(t.n));
            } else {
                plans.push_back(plan_tile(t, autotune && !fm));
                const TileLaunch& g = plans.back();
                last_plan += std::string(t.kind) + (t.ss.iboxes ? " (interval boxes)" : "") +
                             ": lines=" + std::to_string(g.tile_lines) +
                             " block=" + std::to_string(g.block) + " slices=" + std::to_string(g.slices) +
                             " windows=" + std::to_string(g.windows) + " bins=" + (g.fixed ? "fixed" : "float") +
                             " list=" + std::to_string(g.list_cap) + (g.list_packed ? " packed" : "") + "\n";
            }
        }
        const int frame_radix = plan_radix(autotune && !fm);
        last_plan += "convolve: radix=" + std::to_string(frame_radix) + "\n";

        // The bins are zeroed first unless a single launch writes every bin with plain stores.
        const bool single = targets.size() == 1;
        const bool need_zero = !single || (direct_mode ? direct_plans[0].slices > 1 : plans[0].slices > 1);
        if (need_zero) {
            Timed t(this, "kernel_memset_ms");
            hip_check(launch_fill_zero(bins, bin_floats, stream), "fill kernel");
            t.stop();
        } else if (store_details) {
            debug["kernel_memset_ms"].push_back(0.0);  // every bin is stored by the projection
        }
        stats_launches = 0;
        if (count_pairs && !direct_mode && !targets.empty()) {
            d_stats.ensure(kStatsPerLaunch * sizeof(unsigned long long) * targets.size());
            hip_check(launch_fill_zero(d_stats.as<float>(), 2 * kStatsPerLaunch * targets.size(), stream), "fill kernel");
            stats_launches = targets.size();
        }
        for (size_t i = 0; i < targets.size(); i++) {
            const ProjTarget& t = targets[i];
            const bool is_spline = std::strcmp(t.kind, "fixed") != 0;
            const char* tkey = is_spline ? "spline_projection_kernel_ms" : "fixed_projection_kernel_ms";
            if (direct_mode) {
                Timed tm(this, tkey);
                ProjLaunch cfg = direct_plans[i];
                cfg.atomic = (cfg.slices > 1 || !single) ? 1 : 0;
                if (std::strcmp(t.kind, "spline") == 0)
                    hip_check(launch_project_spline(t.c, t.ss, bins, cplx, arc, phase_delay, lut_profile, cfg, stream),
                              "spline projection kernel");
                else
                    hip_check(launch_project_fixed(t.c, t.fs, bins, cplx, arc, phase_delay, lut_profile, cfg, stream),
                              is_spline ? "spline projection kernel" : "fixed projection kernel");
                tm.stop();
            } else {
                TileLaunch cfg = plans[i];
                cfg.atomic |= single ? 0 : 1;
                if (stats_launches) cfg.stats = d_stats.as<unsigned long long>() + kStatsPerLaunch * i;
                const int blocks = tile_blocks(cfg);
                if (!fm && tile_order && blocks >= kMinOrderedBlocks && blocks <= kMaxOrderedTiles) {
                    int* wo = order_buffer(i, blocks);
                    cfg.work = wo;
                    cfg.order = wo + blocks;
                    hip_check(launch_order_tiles(cfg.work, const_cast<int*>(cfg.order), blocks, stream),
                              "tile order kernel");
                }
                Timed tm(this, tkey);  // the projection kernel alone
                launch_target(t, cfg, bins);
                tm.stop();
            }
        }
        ConvParams p = conv_params(d_out);
        p.bins = bins;
        p.radix = frame_radix;
        {
            Timed t(this, "kernel_convolve_demodulate_ms", "kernel_forward_fft_ms");
            hip_check(launch_convolve_demod(p, stream), "convolve/demodulate kernel");
            t.stop();
        }
    }

    // Arguments of the convolution of this frame (prepare() done); advances the
    // noise stream's frame counter.
    ConvParams conv_params(float2* d_out) {
        ConvParams p{};
        p.bins = d_bins.p;
        p.complex_bins = phase_delay ? 1 : 0;
        p.num_lines = (int)lines.size();
        p.num_rf = num_rf;
        p.fft_len = fft_len;
        p.log2_len = log2_len;
        p.twiddle_bits = log2_len / 2;
        p.twiddles = d_tw.as<float2>();
        p.H = d_H.as<float2>();
        p.demod = d_demod.as<float2>();
        p.center = center;
        p.decimation = decimation;
        p.n_out = n_out;
        // fold the inverse transform by the largest power of two dividing the
        // decimation, keeping at least 16 points
        int f = 0;
        while (((decimation >> f) & 1) == 0 && log2_len - (f + 1) >= 4) f++;
        p.fold_log2 = f;
        p.out_stride = decimation >> f;
        p.iq = d_out;
        p.noise_amplitude = noise_amplitude > 0.0f ? noise_amplitude : 0.0f;
        p.noise_seed = noise_seed;
        p.noise_counter = noise_counter++;
        p.noise_counter_ptr = nullptr;
        p.radix = kDefaultRadix;
        return p;
    }

    void host_basis_line0(const SplineDataset& sp, SplineLineBasis& B) {
        const float t = lines[0].t;
        const int mu = knot_interval(sp.knots, t);
        std::memset(&B, 0, sizeof(B));
        B.lo = mu - sp.degree;
        B.count = sp.degree + 1;
        for (int i = 0; i < B.count; i++) B.b[i] = bspline_basis(B.lo + i, sp.degree, t, sp.knots);
    }

    void simulate_device(void* d_out, size_t capacity, int* nl, int* ns) {
        prepare();
        const size_t need = lines.size() * (size_t)n_out;
        if (capacity < need) fail(BCSIM_ERR_CAPACITY, "output capacity too small");
        if (nl) *nl = (int)lines.size();
        if (ns) *ns = n_out;
        last_frame_noalloc = false;
        enqueue_frame(static_cast<float2*>(d_out));
        if (store_details) debug["kernel_memcpy_ms"].push_back(0.0);  // the IQ stays in device memory
        if (store_details || stats_launches) collect_timings();
    }

    bool last_frame_noalloc = false;  // the last frame's bins were in the shared scratch

    // simulate_device without allocating: the bins (and rendered spline points) of
    // the frame live in the library's fixed scratch array of this GPU, which every
    // allocation-free frame on the GPU shares.  Everything else a frame reads must
    // already exist: the simulator's tables are built by its first frame through
    // simulate_device or simulate_host, and a configuration change needs such a
    // frame again.  No launch-shape tuning, block ordering, kernel timing or counters.
    void simulate_device_noalloc(void* d_out, size_t capacity, int* nl, int* ns) {
        require_configured();
        if (lines.empty()) fail(BCSIM_ERR_INVALID_STATE, "No scanlines in scansequence");
        if (profile < 0) fail(BCSIM_ERR_INVALID_STATE, "No beam profile is configured");
        const bool lut_ready = profile != kProfileLut || weight_cutoff <= 0.0f ||
                               (lut_boxes && lut_boxes_cut == weight_cutoff);
        bool basis_ready = true;
        if (!all_equal)
            for (auto& sc : spline_scan) basis_ready = basis_ready && sc->basis.bytes >= lines.size() * sizeof(SplineLineBasis);
        if (!stream || dirty_scan || dirty_conv || dirty_basis || !lut_ready || !basis_ready || fft_len != tables_fft ||
            n_out != tables_out || d_lines.bytes < lines.size() * sizeof(LineGeom))
            fail(BCSIM_ERR_INVALID_STATE,
                 "allocation-free frame: the simulator's tables are not built for its current configuration; "
                 "run one frame through bcsim_simulate_lines_device or bcsim_simulate_lines first");
        const size_t need = lines.size() * (size_t)n_out;
        if (capacity < need) fail(BCSIM_ERR_CAPACITY, "output capacity too small");
        const size_t bytes = noalloc_bytes();
        if (bytes > kNoAllocScratchBytes)
            fail(BCSIM_ERR_CAPACITY, "allocation-free frame needs " + std::to_string(bytes) +
                                         " bytes of scratch (bins and rendered points), more than its fixed " +
                                         std::to_string(kNoAllocScratchBytes) + "; use bcsim_simulate_lines_device");
        hip_check(hipSetDevice(device), "hipSetDevice");
        void* base = nullptr;
        hip_check(noalloc_scratch(device, &base), "hipGetSymbolAddress(scratch)");
        FrameMem fm;
        fm.scratch = static_cast<unsigned char*>(base);
        fm.cap = kNoAllocScratchBytes;
        fm.bins = static_cast<float*>(fm.carve(lines.size() * (size_t)num_rf * (phase_delay ? 8 : 4)));
        const bool details = store_details, counters = count_pairs;
        store_details = count_pairs = false;  // timing events and counters need allocations of their own
        try {
            enqueue_frame(static_cast<float2*>(d_out), &fm);
        } catch (...) {
            store_details = details;
            count_pairs = counters;
            throw;
        }
        store_details = details;
        count_pairs = counters;
        last_frame_noalloc = true;
        if (nl) *nl = (int)lines.size();
        if (ns) *ns = n_out;
    }

    void simulate_host(float* out, size_t capacity, int* nl, int* ns) {
        prepare();
        last_frame_noalloc = false;
        const size_t need = lines.size() * (size_t)n_out;
        if (capacity < need) fail(BCSIM_ERR_CAPACITY, "output capacity too small");
        d_iq.ensure(need * sizeof(float2));
        enqueue_frame(d_iq.as<float2>());
        const size_t bytes = need * sizeof(float2);
        if (bytes >= kStagedCopyMin && host_pageable(out)) {
            copy_staged(out, d_iq.p, bytes);
        } else {
            {
                Timed t(this, "kernel_memcpy_ms");
                hip_check(hipMemcpyAsync(out, d_iq.p, bytes, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
                t.stop();
            }
            hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        }
        if (store_details || stats_launches) collect_timings();
        if (nl) *nl = (int)lines.size();
        if (ns) *ns = n_out;
    }

    // Frames copied into pageable host memory pass through two pinned pieces of
    // kStagePiece bytes: the DMA of piece c + 1 runs while host threads move piece
    // c into the caller's buffer (the runtime's own path to pageable memory is
    // about a third of the DMA rate).
    static constexpr size_t kStagedCopyMin = (size_t)8 << 20;
    static constexpr size_t kStagePiece = (size_t)32 << 20;
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};

    void copy_staged(void* dst, const void* src, size_t bytes) {
        for (int i = 0; i < 2; i++) {
            if (!stage[i]) hip_check(hipHostMalloc(&stage[i], kStagePiece, 0), "hipHostMalloc");
            if (!stage_ev[i]) hip_check(hipEventCreateWithFlags(&stage_ev[i], hipEventDisableTiming), "hipEventCreate");
        }
        const size_t pieces = (bytes + kStagePiece - 1) / kStagePiece;
        auto piece_len = [&](size_t c) { return std::min(kStagePiece, bytes - c * kStagePiece); };
        auto issue = [&](size_t c) {
            hip_check(hipMemcpyAsync(stage[c % 2], (const char*)src + c * kStagePiece, piece_len(c),
                                     hipMemcpyDeviceToHost, stream),
                      "hipMemcpyAsync");
            hip_check(hipEventRecord(stage_ev[c % 2], stream), "hipEventRecord");
        };
        Timed t(this, "kernel_memcpy_ms");
        issue(0);
        for (size_t c = 0; c < pieces; c++) {
            if (c + 1 < pieces) issue(c + 1);  // its pinned piece was emptied in the previous round
            if (c + 1 == pieces) t.stop();
            hip_check(hipEventSynchronize(stage_ev[c % 2]), "hipEventSynchronize");
            parallel_copy((char*)dst + c * kStagePiece, stage[c % 2], piece_len(c));
        }
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void debug_projection(float* out, size_t capacity, int* nl, int* nrf) {
        if (!d_bins.p || dirty_scan) fail(BCSIM_ERR_INVALID_STATE, "no simulated frame");
        if (last_frame_noalloc)
            fail(BCSIM_ERR_INVALID_STATE, "the last frame ran without allocation: its bins were in the shared scratch");
        const size_t L = lines.size();
        if (nl) *nl = (int)L;
        if (nrf) *nrf = num_rf;
        if (!out) return;  // shape query
        if (capacity < L * (size_t)num_rf) fail(BCSIM_ERR_CAPACITY, "output capacity too small");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (phase_delay) {
            hip_check(hipMemcpy(out, d_bins.p, L * num_rf * sizeof(float2), hipMemcpyDeviceToHost), "hipMemcpy");
        } else {
            std::vector<float> re(L * (size_t)num_rf);
            hip_check(hipMemcpy(re.data(), d_bins.p, re.size() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
            for (size_t i = 0; i < re.size(); i++) {
                out[2 * i] = re[i];
                out[2 * i + 1] = 0.0f;
            }
        }
        if (nl) *nl = (int)L;
        if (nrf) *nrf = num_rf;
    }
};

// Run a C-ABI body: map exceptions to a status code and this thread's message.
template <typename F>
bcsim_status guarded(F&& f) {
    try {
        f();
        bcsim_hip::last_error().clear();
        return BCSIM_OK;
    } catch (const bcsim_hip::BcsimError& e) {
        bcsim_hip::last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        bcsim_hip::last_error() = "host allocation failed";
        return BCSIM_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        bcsim_hip::last_error() = e.what();
        return BCSIM_ERR_INTERNAL;
    }
}

inline void need(bcsim_sim* s) {
    if (!s) bcsim_hip::fail(BCSIM_ERR_INVALID_ARGUMENT, "null simulator handle");
}
