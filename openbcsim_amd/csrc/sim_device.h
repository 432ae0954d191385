// Device-visible data layout of the simulation path and the launch interface
// between the host orchestration (sim_host.cpp) and the kernels (sim_kernels.hip).
//
// HBM layout (per simulator handle):
//   lines        LineGeom[L]                       64 B per line
//   fixed sets   float4[N] (x, y, z, amplitude)    16 B per scatterer, one load
//   spline sets  float  cx[M][N], cy[M][N], cz[M][N], amp[N]
//                (control-point-major, so the p+1 active control points of a
//                 line are p+1 coalesced streams; the reference's layout idea,
//                 GpuScatterers.cpp:257)
//   spline basis SplineLineBasis[L] per spline set  (host-evaluated, float)
//   bins         float[L][num_rf] (real) or float2[L][num_rf] (phase delay)
//   H            float2[N_fft]   excitation spectrum * Hilbert mask / N_fft
//   demod        float2[n_out]   exp(-2 pi i nfd j D)
//   twiddles     float2[2*T]     two-level table of exp(-2 pi i k / N_fft)
//   iq           float2[L][n_out]
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace bcsim_hip {

constexpr int kMaxSplineDegree = 4;   // common_definitions.h:2
constexpr int kMaxFftLen = 16384;     // one line in 128 KiB of LDS
constexpr int kDefaultRadix = 8;       // FFT pass radix when not set or tuned (8 measured faster at every length)
constexpr int kMinFftLen = 16;
constexpr int kLdsBytes = 160 * 1024; // gfx950 LDS per workgroup
// Default weight cutoff (parameter weight_cutoff): profile weights below 1e-8 of
// the peak are dropped, 6.07 sigma for a Gaussian.  A dropped pair changes its
// bin by less than 1e-8 of a full-weight echo of the same amplitude, below the
// rounding of a float32 sum (6e-8 relative).
constexpr float kDefaultWeightCutoff = 1e-8f;

struct alignas(16) LineGeom {
    float ox, oy, oz, t;
    float rx, ry, rz, p0;
    float lx, ly, lz, p1;
    float ex, ey, ez, p2;
};

struct alignas(16) SplineLineBasis {
    int lo;                          // first control point index of the sum
    int count;                       // number of terms (degree + 1, or M with sum_all_cs)
    float b[kMaxSplineDegree + 1];   // basis values for lo .. lo+count-1
    float pad;
};

enum ProfileKind { kProfileGaussian = 0, kProfileLut = 1 };

struct ProfileParams {
    // Gaussian: exp(-(l*l*inv_tsl + e*e*inv_tse)), inv_ts* = 1/(2 sigma^2); a pair whose
    // exponent exceeds g_cut (-ln of the weight cutoff; +inf without one) weighs 0
    float inv_tsl, inv_tse;
    float g_cut;
    // LUT (BeamProfile.cpp:79-130): t = (x - x_min) * inv_dx, trilinear, zero outside;
    // read as quad pairs (see lut_weight): quad[2*((r*ne + e)*nl + l) + s] holds slice r + s
    const float4* quad;
    int nr, nl, ne;
    float r_min, l_min, e_min;
    float inv_dr, inv_dl, inv_de;
    // Weight cutoff (null / 0 without one): per radial record index r, the
    // inclusive ranges (l_lo, l_hi, e_lo, e_hi) of the records (r, l, e) holding a
    // sample of magnitude >= w_cut; a pair whose record lies outside weighs 0 (its
    // weight, a convex combination of smaller samples, is below w_cut), and so
    // does a pair whose interpolated |weight| < w_cut.  Level k of the sparse table
    // slice_box + k*nr holds the union over records [r, r + 2^k).
    const int4* slice_box;
    float w_cut;
};

struct ProjCommon {
    const LineGeom* lines;
    int num_lines;
    int num_rf;
    // radial bin = floor(r * scale / denom + 0.5) in double (CpuAlgorithm.cpp:75, :161-162);
    // evaluated as r*scale*inv_denom with an exact-division fallback near half-samples.
    double bin_scale;
    double bin_denom;
    double bin_inv_denom;
    // sub-sample phase: phase = (bin - true_index) * phase_per_sample, true_index = r*scale/denom
    float phase_per_sample;
    ProfileParams prof;
    // Fixed-point bins (TileLaunch::fixed): a contribution w*a enters the LDS bins
    // as the 64-bit integer rint(w * a * amp_scale), amp_scale = 2^k with every
    // |w*a| * amp_scale below 2^51 and its sum over the launch's scatterers below
    // 2^62, so no bin sum can overflow; a bin's sum is read back as
    // float(sum * inv_amp_scale).
    float amp_scale;
    float inv_amp_scale;
};

struct FixedSet {
    const float4* pts;
    int n;
};

struct ChunkBox;  // (below)

struct SplineSet {
    const float* cx;
    const float* cy;
    const float* cz;
    const float* amp;
    const SplineLineBasis* basis;   // [num_lines]
    int n;
    int num_cp;
    // or null: per chunk and knot interval, the box of the chunk's active control
    // points on that interval ([nchunks][n_iv], n_iv = num_cp - degree); a spline
    // point on a line lies in the box of its line's interval (the basis values are
    // non-negative and sum to one), which is tighter than the box of all its points
    const ChunkBox* iboxes;
    int n_iv;
};

struct ProjLaunch {
    int tile_lines;    // lines held in LDS by one workgroup
    int slices;        // scatterer slices (grid.y); >1 flushes with global atomics
    int block;         // threads per workgroup
    int atomic;        // flush with global atomics (several slices or several datasets)
};

struct ConvParams {
    const void* bins;       // float[L][num_rf] or float2[L][num_rf]
    int complex_bins;
    int num_lines;
    int num_rf;
    int fft_len;
    int log2_len;
    int twiddle_bits;       // fine table holds 2^twiddle_bits entries
    const float2* twiddles; // coarse[2^(log2_len - twiddle_bits)] then fine[2^twiddle_bits]
    // H[k] = mask[k] * FFT_N(excitation)[k] / N * exp(2 pi i k center / N) for k <= N/2,
    // zero above: the spectrum of the delay-compensated output x[center + n]
    const float2* H;        // [fft_len]
    const float2* demod;    // [n_out]
    int center;
    int decimation;
    int n_out;
    // the inverse transform is N / 2^fold_log2 points long (2^fold_log2 = the largest
    // power of two dividing the decimation); output m is its sample out_stride * m
    int fold_log2;
    int out_stride;
    float2* iq;             // [L][n_out]
    // additive noise on the bins before convolution (CpuAlgorithm.cpp:340-347)
    float noise_amplitude;
    uint32_t noise_seed;
    uint32_t noise_counter;
    // when set, the frame counter is read from device memory instead (batched
    // frames: launch_fill_segments advances it at the start of every frame)
    const uint32_t* noise_counter_ptr;
    int radix;              // largest FFT pass radix: 4 or 8
};

// LDS points (float2) of one line's convolution buffer; its twiddle table follows
// it.  Real bins without noise, folded (fold_log2 >= 1): the N/2-point transform
// of the packed samples and the folded inverse both fit N/2 points.
// LDS slot of FFT point i: one float2 of padding after every 2^kConvPadShift points,
// so that the early Stockham passes, whose lanes write 2^k points apart, spread
// their writes over the banks.
constexpr int kConvPadShift = 3;
__host__ __device__ inline int conv_slot(int i) { return i + (i >> kConvPadShift); }

// LDS float2 slots of the convolution's transform buffer (padded).
__host__ __device__ inline int conv_buffer_len(const ConvParams& p) {
    const bool real_in = !p.complex_bins && !(p.noise_amplitude > 0.0f);
    return conv_slot((real_in && p.fold_log2 > 0) ? p.fft_len / 2 : p.fft_len);
}

// ---- culled tile projection (project_tile.hip) ----

constexpr int kChunk = 64;      // scatterers per chunk: one per lane of a wave
constexpr int kTileLineLds = 96; // LDS bytes per tile line beside its bins: LineGeom + SplineLineBasis
constexpr int kMaxListCap = 2048;  // chunk list entries held in LDS between drains

// Axis-aligned box around the (finite) positions a chunk can take: the points
// of a fixed chunk, or every control point of a spline chunk (a B-spline stays
// in the convex hull of its control points).  hx < 0 marks an empty chunk.
struct alignas(16) ChunkBox {
    float cx, cy, cz, pad0;
    float hx, hy, hz, pad1;
};

// Support of the beam-profile weight in beam-local coordinates.  Outside it the
// weight is exactly zero in float (LUT: outside the table, BeamProfile.cpp:97-105;
// Gaussian: exponent beyond float's underflow), so skipping those pairs is exact.
struct CullParams {
    float r_per_bin;          // metres of radial distance per bin (denom / scale)
    float r_lo, r_hi;         // radial: bin window [0, num_rf) and LUT r support
    float l_lo, l_hi;         // LUT lateral support
    float e_lo, e_hi;         // LUT elevational support
    float g_a, g_b, g_max;    // Gaussian: a*l^2 + b*e^2 < g_max
    float lut_dl, lut_de;     // LUT grid spacing, 0 for a degenerate table (weight cutoff: records to metres)
};

struct TileLaunch {
    int tile_lines;           // lines whose bins one workgroup keeps in LDS (<= 32)
    int windows;              // radial windows per line; workgroup = tile * windows + window
    int window_bins;          // bins per window (the last window may be shorter)
    int slices;               // chunk ranges per (tile, window) (grid.y)
    int chunks_per_slice;
    int block;                // threads per workgroup (multiple of 64)
    int atomic;               // flush with global atomics
    int fixed;                // LDS bins as 64-bit fixed point (ProjCommon::amp_scale), else float
    int list_cap;             // chunk list entries in LDS (multiple of 64)
    int list_packed;          // 1: an entry is one 32-bit word, chunk << 4 | line mask (tiles of at most
                              // 4 lines, fewer than 2^28 chunks); 0: two words, chunk and mask
    const int* order;         // workgroup w runs block order[w] (heaviest first), or null: block w
    int* work;                // per-block hit counts added up by the launch, or null
    const int* chunks;        // chunk visiting order (a permutation of [0, nchunks)), or null: index order
    // or null: for lines from a common origin, two ascending bounds per position p of the visiting
    // order, far[p] (largest centre distance + box half-diagonal over positions <= p) then near[p]
    // (smallest centre distance - half-diagonal over positions >= p); a radial window visits only
    // the positions whose chunks can reach its distances
    const float* chunk_dist;
    // or null: per launch, [0] profile-weight evaluations (LUT record gathers /
    // Gaussian exponentials), [1] LDS bin additions and [2] (chunk, line) pairs
    // the cull kept (each evaluated by a wave of 64 lanes), added up by the launch
    unsigned long long* stats;
};

// Blocks one order_tiles launch orders (one workgroup walks them three times).
constexpr int kMaxOrderedTiles = 65536;

// One projection job of a batch: a (simulator, dataset) pair with everything its
// tile kernel would receive as arguments.
enum JobFlags { kJobArc = 1, kJobPhase = 2, kJobLut = 4, kJobSpline = 8, kJobShared = 16, kJobFixed = 32 };

struct ProjJob {
    ProjCommon c;
    CullParams cp;
    FixedSet fs;
    SplineSet ss;
    const ChunkBox* boxes;
    float* bins;
    TileLaunch g;
    int nchunks;
    int flags;
    int blocks_per_slice;     // tiles * windows
};

// A region of floats to zero before a batched frame.
struct FillSeg {
    float* p;
    size_t n;
};

// The blocks of one batched projection job to order by their work counts:
// work[n] and the job's n * slices items.
struct OrderSeg {
    int* work;
    int2* items;
    int n;
    int slices;
};

// Scratch of the allocation-free frame (bcsim_simulate_lines_device_noalloc): a
// fixed device array of this many bytes per GPU in sim_kernels.hip, holding one
// frame's projection bins and rendered spline points.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
constexpr int kMaxScratchDevices = 64;

// Device address of the scratch array on `device` (the current device), looked
// up once per device with hipGetSymbolAddress and kept.
hipError_t noalloc_scratch(int device, void** ptr);

// Let kernel `func` use up to `bytes` of dynamic LDS: the attribute is raised to
// the largest dynamic LDS any launch of this library asks for, once per kernel
// and device (sim_kernels.hip), so launches do not pay for the call.
hipError_t allow_lds(const void* func, size_t bytes);

// Launchers (sim_kernels.hip).  All enqueue on `stream` and return the launch status.
hipError_t launch_fill_zero(float* p, size_t n_floats, hipStream_t stream);
hipError_t launch_project_fixed(const ProjCommon& c, const FixedSet& s, void* bins, bool complex_bins,
                                bool arc, bool phase, bool lut, const ProjLaunch& cfg,
                                hipStream_t stream);
hipError_t launch_project_spline(const ProjCommon& c, const SplineSet& s, void* bins, bool complex_bins,
                                 bool arc, bool phase, bool lut, const ProjLaunch& cfg,
                                 hipStream_t stream);
hipError_t launch_render_spline(const SplineSet& s, const SplineLineBasis& basis, float4* out,
                                hipStream_t stream);
hipError_t launch_convolve_demod(const ConvParams& p, hipStream_t stream);
int conv_block_threads(int fft_len);

// project_tile.hip: fixed (spline == nullptr) or spline dataset, chunked and
// boxed by the host; `shared_origin` when every line starts at lines[0]'s origin.
hipError_t launch_project_tile(const ProjCommon& c, const CullParams& cp, const FixedSet* fixed,
                               const SplineSet* spline, const ChunkBox* boxes, int nchunks, void* bins,
                               bool arc, bool phase, bool lut, bool shared_origin, const TileLaunch& cfg,
                               hipStream_t stream);
// Bytes of one LDS bin of the tile kernel: float or 64-bit fixed point, real or complex.
inline size_t tile_bin_bytes(bool complex_bins, bool fixed) { return (fixed ? 8 : 4) * (complex_bins ? 2 : 1); }
size_t tile_lds_bytes(int window_bins, size_t bin_bytes, int tile_lines, int list_cap, int list_packed);
// Orders blocks 0..n-1 (n <= kMaxOrderedTiles) by work[], heaviest first, into
// order[] and clears work[] for the next tile launch to fill.
hipError_t launch_order_tiles(int* work, int* order, int n, hipStream_t stream);
// launch_order_tiles for several jobs at once: one workgroup per segment.
hipError_t launch_order_segments(const OrderSeg* segs, int n_segs, hipStream_t stream);

// The JobFlags that select projection code: arc, phase, LUT, spline, and a
// shared origin (which only matters for fixed points with arc projection).
inline int batch_variant(int flags) {
    int f = flags & (kJobArc | kJobPhase | kJobLut | kJobSpline | kJobShared | kJobFixed);
    if (!(f & kJobArc) || (f & kJobSpline)) f &= ~kJobShared;
    return f;
}

// LDS bytes of one line's convolution: the padded transform buffer and the
// two-level twiddle table.
__host__ __device__ inline size_t conv_lds_bytes(const ConvParams& p) {
    return ((size_t)conv_buffer_len(p) + (size_t)(p.fft_len >> p.twiddle_bits) + ((size_t)1 << p.twiddle_bits)) *
           sizeof(float2);
}

// Kernel variant of a batched convolution job: workgroup size class (1024 threads
// for FFTs of 4096 points and more, else 256), radix and bin kind.
enum ConvVariant { kConvComplex = 1, kConvRadix8 = 2, kConvWide = 4 };
inline int conv_variant(const ConvParams& p, int block_threads) {
    return (block_threads == 1024 ? kConvWide : 0) | (p.radix == 8 ? kConvRadix8 : 0) |
           (p.complex_bins ? kConvComplex : 0);
}

// Batched frames.  launch_project_batch: persistent workgroups take the items of
// one kernel variant (batch_variant of every item's job) from `counter` (reset by
// launch_fill_segments, which also adds one to each of the n_noise per-job noise
// frame counters).  launch_convolve_batch: one workgroup per item of one
// convolution variant (conv_variant), `lds` bytes each.
hipError_t launch_project_batch(int variant, const ProjJob* jobs, const int2* items, int n_items, int* counter,
                                size_t lds, int workgroups, hipStream_t stream);
hipError_t launch_convolve_batch(int variant, const ConvParams* jobs, const int2* items, int n_items, size_t lds,
                                 hipStream_t stream);
hipError_t launch_fill_segments(const FillSeg* segs, int n_segs, int* counters, int n_counters,
                                uint32_t* noise_counters, int n_noise, hipStream_t stream);

}  // namespace bcsim_hip
