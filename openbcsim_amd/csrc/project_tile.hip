// Culled tile projection for gfx950: the production scatterer-projection kernel.
//
// Work decomposition.  A workgroup owns a tile of `tile_lines` consecutive scan
// lines (at most 32) and one radial window of `window_bins` bins of each of
// them; the bins of that (tile, window) block live in LDS for the whole kernel.
// Workgroup w = tile * windows + window, and the hardware deals workgroups to
// the 8 XCDs round-robin, so with 8 windows each XCD handles one radial window
// of every tile and touches only an eighth of the beam-profile table (measured:
// L2 misses fall 20x on the 1M-scatterer volume).  That workload is bound by the
// vector-memory path of the table gathers rather than by L2 misses, and one
// window (whole lines in LDS) is the faster default; windows exist for lines
// whose bins do not fit in LDS and for the tuner.  grid.y optionally splits the
// chunks into slices.
//
// The host stores every dataset in Morton order of position and keeps one
// axis-aligned box per chunk of 64 scatterers (ChunkBox).  Filling a chunk list
// in LDS a step at a time and draining it when full, a workgroup
//   (A) tests every chunk box -- first against the radial window, then against
//       each line's beam support (outside which the weight is exactly zero in
//       float: LUT table extent, Gaussian underflow) -- and appends survivors
//       with a per-line hit mask to a list in LDS;
//   (B) hands listed chunks to its waves, one chunk per wave and one scatterer
//       per lane; each lane projects its scatterer onto every hit line and adds
//       into the LDS bins.
// The block leaves LDS once: plain stores when the workgroup saw every chunk,
// otherwise atomic adds of the non-zero bins.
//
// LDS bins are float (ds_add_f32) or 64-bit fixed point (ds_add_u64, FIXED).
// On gfx950 an LDS float atomic add retires about one lane per three cycles
// while an integer one retires a wave's 64 lanes about ten times faster
// (measured: ~2e11 against ~2e12 lane additions per second on the chip), so the
// fixed-point bins take the additions off the kernel's critical unit at twice
// the LDS per bin.  Each contribution is rounded once to the fixed-point grid
// (ProjCommon::amp_scale; a quantum of about 2^-61 of the largest possible bin
// sum) and the integer sums are exact, so a block's bins do not depend on the
// order of its additions.
//
// Skipped pairs would have added w*a == 0 to a bin, so the sums equal those of
// the direct evaluation up to float summation order (fixed point: up to the
// rounding of each contribution to the grid).  When every line starts at
// the same origin (sector and pyramidal scans) the arc-projected distance
// |p - o|, its bin and the radial part of the LUT lookup are computed once per
// scatterer instead of once per line.
//
// Compiled with -ffp-contract=off (see proj_math.h).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "proj_math.h"
#include "sim_device.h"

namespace bcsim_hip {

// Can a point of box B lie inside line L's lateral and elevational support
// (LUT extent; for the Gaussian, where expf is not yet zero)?  dx, dy, dz: box
// centre minus line origin; m: the margin.
template <bool LUT>
__device__ __forceinline__ bool lateral_hits(const CullParams& cp, const LineGeom& L, const ChunkBox& B, float dx,
                                             float dy, float dz, float m) {
    const float lc = dx * L.lx + dy * L.ly + dz * L.lz;
    const float hl = B.hx * fabsf(L.lx) + B.hy * fabsf(L.ly) + B.hz * fabsf(L.lz) + m;
    const float ec = dx * L.ex + dy * L.ey + dz * L.ez;
    const float he = B.hx * fabsf(L.ex) + B.hy * fabsf(L.ey) + B.hz * fabsf(L.ez) + m;
    if (LUT) return (lc + hl > cp.l_lo) && (lc - hl < cp.l_hi) && (ec + he > cp.e_lo) && (ec - he < cp.e_hi);
    const float dl = fmaxf(0.0f, fabsf(lc) - hl), de = fmaxf(0.0f, fabsf(ec) - he);
    return !(cp.g_a * dl * dl + cp.g_b * de * de >= cp.g_max);
}

// Can a point of box B have a non-zero weight and a bin inside [r_lo, r_hi) on
// line L?  Conservative: bounds are padded by a relative margin that covers the
// float rounding of the per-point evaluation.
template <bool ARC, bool LUT>
__device__ __forceinline__ bool chunk_hits(const CullParams& cp, float r_lo, float r_hi, const LineGeom& L,
                                           const ChunkBox& B) {
    if (!(B.hx >= 0.0f)) return false;  // empty chunk
    const float dx = B.cx - L.ox, dy = B.cy - L.oy, dz = B.cz - L.oz;
    const float hn = sqrtf(B.hx * B.hx + B.hy * B.hy + B.hz * B.hz);
    const float dn = sqrtf(dx * dx + dy * dy + dz * dz);
    const float m = 1e-6f * (fabsf(B.cx) + fabsf(B.cy) + fabsf(B.cz) + fabsf(L.ox) + fabsf(L.oy) + fabsf(L.oz) +
                             dn + hn) + 1e-30f;
    const float rc = dx * L.rx + dy * L.ry + dz * L.rz;
    const float hr = B.hx * fabsf(L.rx) + B.hy * fabsf(L.ry) + B.hz * fabsf(L.rz) + m;
    bool hit_r;
    if (ARC) {
        // r = +-|p - o| with the sign of the projection onto the beam direction
        const float rmin = fmaxf(0.0f, dn - hn) - m, rmax = dn + hn + m;
        const bool pos = (rc + hr > 0.0f) && (rmax > r_lo) && (rmin < r_hi);
        const bool neg = (rc - hr < 0.0f) && (-rmin > r_lo) && (-rmax < r_hi);
        hit_r = pos || neg;
    } else {
        hit_r = (rc + hr > r_lo) && (rc - hr < r_hi);
    }
    if (!hit_r) return false;
    return lateral_hits<LUT>(cp, L, B, dx, dy, dz, m);
}

// chunk_hits for lines from a common origin with arc projection, the distance
// terms of the box (computed once per chunk by the caller) given: dx, dy, dz the
// box centre from the origin, m the margin, pos_ok / neg_ok whether the box's
// distance range meets the window on the positive / negative side.
template <bool LUT>
__device__ __forceinline__ bool chunk_hits_shared(const CullParams& cp, const LineGeom& L, const ChunkBox& B,
                                                  float dx, float dy, float dz, float m, bool pos_ok,
                                                  bool neg_ok) {
    const float rc = dx * L.rx + dy * L.ry + dz * L.rz;
    const float hr = B.hx * fabsf(L.rx) + B.hy * fabsf(L.ry) + B.hz * fabsf(L.rz) + m;
    if (!((pos_ok && rc + hr > 0.0f) || (neg_ok && rc - hr < 0.0f))) return false;
    return lateral_hits<LUT>(cp, L, B, dx, dy, dz, m);
}

// One LDS bin addition: float, or a contribution v already scaled by amp_scale
// (|v| < 2^51 by the choice of amp_scale), rounded to the nearest integer and
// added as a 64-bit two's-complement value.  The rounding and conversion are one
// double addition: v + 1.5 * 2^52 lies in [2^52, 2^53), where doubles are the
// integers, so the sum is 1.5 * 2^52 + rint(v) exactly, and its bit pattern minus
// that of 1.5 * 2^52 is rint(v) (the low word of the constant is zero).
__device__ __forceinline__ void lds_add(float* b, float v) { atomicAdd(b, v); }
__device__ __forceinline__ void lds_add(long long* b, float v) {
    const double t = (double)v + 6755399441055744.0;  // 1.5 * 2^52
    const long long q = __double_as_longlong(t) - 0x4338000000000000LL;
    atomicAdd(reinterpret_cast<unsigned long long*>(b), (unsigned long long)q);
}

// A bin sum as float.
__device__ __forceinline__ float bin_value(float v, const ProjCommon&) { return v; }
__device__ __forceinline__ float bin_value(long long v, const ProjCommon& c) {
    return (float)((double)v * (double)c.inv_amp_scale);
}

// Weight times amplitude added at window-relative bin lb (CpuAlgorithm.cpp:77-95).
// With fixed-point bins the amplitude was scaled by amp_scale, a power of two, so
// w is the float product w*a times amp_scale exactly.
template <bool PHASE, class Acc>
__device__ __forceinline__ void add_bin(const ProjCommon& c, Acc* lbins, int lb, int bin, double true_index,
                                        float w, Tally& t) {
    if (w == 0.0f) return;  // adding +-0 leaves the bin unchanged
    t.adds += PHASE ? 2 : 1;
    if (PHASE) {
        const float ph = (float)((double)bin - true_index) * c.phase_per_sample;
        float s, co;
        sincosf(ph, &s, &co);
        lds_add(&lbins[2 * lb], w * co);
        lds_add(&lbins[2 * lb + 1], w * s);
    } else {
        lds_add(&lbins[lb], w);
    }
}

// add_bin with the sub-sample rotation (cos, sin) of the phase already computed
// (the same for every line when the lines share an origin).
template <bool PHASE, class Acc>
__device__ __forceinline__ void add_bin_rot(Acc* lbins, int lb, float w, float co, float s, Tally& t) {
    if (w == 0.0f) return;
    t.adds += PHASE ? 2 : 1;
    if (PHASE) {
        lds_add(&lbins[2 * lb], w * co);
        lds_add(&lbins[2 * lb + 1], w * s);
    } else {
        lds_add(&lbins[lb], w);
    }
}

// Directions of one line, held in registers.
struct LineDirs {
    float rx, ry, rz, lx, ly, lz, ex, ey, ez;
};

__device__ __forceinline__ LineDirs line_dirs(const LineGeom& L) {
    return LineDirs{L.rx, L.ry, L.rz, L.lx, L.ly, L.lz, L.ex, L.ey, L.ez};
}

template <bool PHASE, bool LUT, class Acc>
__device__ __forceinline__ void emit(const ProjCommon& c, Acc* lbins, int b0, int nb, float r, float l, float e,
                                     int bin, double ti, float amp, Tally& t) {
    const int lb = bin - b0;
    if (bin < 0 || lb < 0 || lb >= nb) return;
    const float w = (LUT ? lut_weight(c.prof, r, l, e, t) : gauss_weight(c.prof, l, e, t)) * amp;
    add_bin<PHASE>(c, lbins, lb, bin, ti, w, t);
}

// With a LUT weight cutoff: narrow the lateral and elevational support of the
// cull to the records of slices [s0, s1] that hold samples above the cutoff
// (the union of their boxes, from the sparse table), in metres with a record of
// slack below and two above.  Returns false when none of those slices has such
// a record, so no pair of the chunk can contribute.
__device__ __forceinline__ bool narrow_support(const ProfileParams& P, CullParams& cp, int s0, int s1) {
    const int n = s1 - s0 + 1;
    const int k = 31 - __builtin_clz(n);  // largest power of two not above n
    const int4 a = P.slice_box[k * P.nr + s0];
    const int4 b = P.slice_box[k * P.nr + s1 - (1 << k) + 1];
    const int l_lo = min(a.x, b.x), l_hi = max(a.y, b.y), e_lo = min(a.z, b.z), e_hi = max(a.w, b.w);
    if (l_lo > l_hi || e_lo > e_hi) return false;
    cp.l_lo = fmaxf(cp.l_lo, P.l_min + (float)(l_lo - 1) * cp.lut_dl);
    cp.l_hi = fminf(cp.l_hi, P.l_min + (float)(l_hi + 2) * cp.lut_dl);
    cp.e_lo = fmaxf(cp.e_lo, P.e_min + (float)(e_lo - 1) * cp.lut_de);
    cp.e_hi = fminf(cp.e_hi, P.e_min + (float)(e_hi + 2) * cp.lut_de);
    return true;
}

// One (tile, window, slice) block of work: everything a workgroup of the tile
// kernel does.  `wg` = tile * windows + window; `lds` is the dynamic LDS and
// `list_n` a workgroup-shared counter.  FIXED: 64-bit fixed-point LDS bins.
// CP_AHEAD: a spline's active control points are all loaded before the first is
// used (otherwise one control point at a time).
template <bool ARC, bool PHASE, bool LUT, bool SPLINE, bool SHARED, bool FIXED, bool CP_AHEAD = true>
__device__ __forceinline__ void tile_block(const ProjCommon& c, const CullParams& cp, const FixedSet& fs,
                                           const SplineSet& ss, const ChunkBox* __restrict__ boxes, int nchunks,
                                           float* __restrict__ out, const TileLaunch& g, int wg, int slice,
                                           float* lds, int& list_n) {
    constexpr int E = PHASE ? 2 : 1;
    using Acc = typename std::conditional<FIXED, long long, float>::type;
    const int window = wg % g.windows;
    const int tile = wg / g.windows;
    const int line0 = tile * g.tile_lines;
    const int nl = min(g.tile_lines, c.num_lines - line0);
    const int b0 = window * g.window_bins;
    const int nb = min(g.window_bins, c.num_rf - b0);
    const int stride = g.window_bins * E;  // LDS bins (Acc) per line
    // LDS: the tile's line geometry and spline basis (read for every hit; kept here
    // rather than re-read from global memory), then the bins, then the chunk list
    LineGeom* tl = reinterpret_cast<LineGeom*>(lds);
    SplineLineBasis* tb = reinterpret_cast<SplineLineBasis*>(tl + g.tile_lines);
    Acc* bins = reinterpret_cast<Acc*>(tb + g.tile_lines);  // 32-byte aligned
    int* list_chunk = reinterpret_cast<int*>(bins + (size_t)g.tile_lines * stride);
    unsigned* list_mask = reinterpret_cast<unsigned*>(list_chunk + g.list_cap);  // two-word entries
    unsigned* list_word = reinterpret_cast<unsigned*>(list_chunk);              // packed entries
    {
        const float4* src = reinterpret_cast<const float4*>(c.lines + line0);
        float4* dst = reinterpret_cast<float4*>(tl);
        for (int i = threadIdx.x; i < nl * 4; i += blockDim.x) dst[i] = src[i];
        if (SPLINE) {
            const float4* bs = reinterpret_cast<const float4*>(ss.basis + line0);
            float4* bd = reinterpret_cast<float4*>(tb);
            for (int i = threadIdx.x; i < nl * 2; i += blockDim.x) bd[i] = bs[i];
        }
    }
    for (int i = threadIdx.x; i < nl * stride; i += blockDim.x) bins[i] = Acc(0);
    if (threadIdx.x == 0) list_n = 0;
    __shared__ unsigned long long wg_stats[3];  // the block's launch statistics (TileLaunch::stats)
    if (g.stats && threadIdx.x < 3) wg_stats[threadIdx.x] = 0;

    // radial extent of this window, intersected with the profile's radial support
    const float r_lo = fmaxf(cp.r_lo, ((float)b0 - 0.5f) * cp.r_per_bin);
    const float r_hi = fminf(cp.r_hi, ((float)(b0 + nb) - 0.5f) * cp.r_per_bin);

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform per wave
    const int nwaves = blockDim.x >> 6;
    int cbeg = slice * g.chunks_per_slice;
    int cend = min(nchunks, cbeg + g.chunks_per_slice);
    if (SHARED && ARC && g.chunk_dist && g.windows > 1) {
        // the positions whose chunks can hold points at distances [r_lo, r_hi]: before the
        // first far bound >= r_lo every chunk ends nearer, from the first near bound > r_hi on
        // every chunk starts farther (a relative 1e-5 of slack for float rounding).  Wave 0
        // finds the first, wave 1 the second, each by a 64-way search: 3 rounds of one load per
        // lane for 2^18 chunks instead of 18 dependent loads.
        __shared__ int win_range[2];
        if (wave < 2) {
            const float* bound = wave == 0 ? g.chunk_dist : g.chunk_dist + nchunks;
            const float x = wave == 0 ? fmaxf(r_lo, 0.0f) * (1.0f - 1e-5f) : r_hi * (1.0f + 1e-5f) + 1e-30f;
            // first position p in [0, nchunks) with bound[p] >= x (wave 0) or > x (wave 1)
            int a = 0, b = nchunks;
            while (a < b) {
                const int stride = (b - a + 63) >> 6;
                const int q = a + lane * stride;
                bool before = false;
                if (q < b) before = wave == 0 ? bound[q] < x : bound[q] <= x;  // a prefix of the lanes
                const int cnt = __popcll(__ballot(before));
                if (cnt == 0) {
                    b = a;
                } else {
                    const int na = a + (cnt - 1) * stride + 1;
                    b = min(b, a + cnt * stride);
                    a = na;
                }
            }
            if (lane == 0) win_range[wave] = a;
        }
        __syncthreads();
        cbeg = max(cbeg, win_range[0]);
        cend = min(cend, win_range[1]);
    }
    const int N = SPLINE ? ss.n : fs.n;

    // The list is filled `step` chunks at a time and drained when another step
    // might not fit, so a workgroup synchronises about twice per step and the
    // drains hand its waves long lists.
    const int step = min((int)blockDim.x, g.list_cap);
    int hits = 0;  // (chunk, line) pairs this thread listed
    Tally tally;
    __syncthreads();
    // Tiles of up to kRegLines lines sharing an origin keep the origin and line
    // directions in registers for the whole kernel: the per-hit loop then reads
    // no LDS besides its additions, so no wait on the LDS queue (which holds the
    // additions of earlier hits) sits between a hit's table gathers and the next.
    constexpr int kRegLines = 4;
    const bool reg_lines = SHARED && ARC && !SPLINE && nl <= kRegLines;
    LineDirs G[kRegLines];
    float org_x = 0.0f, org_y = 0.0f, org_z = 0.0f;
    if (SHARED && ARC && !SPLINE) {
#pragma unroll
        for (int k = 0; k < kRegLines; k++) G[k] = line_dirs(tl[k < nl ? k : 0]);
        org_x = tl[0].ox;
        org_y = tl[0].oy;
        org_z = tl[0].oz;
    }
    for (int next = cbeg; next < cend;) {
        // (A) cull the next `step` chunks
        const int pos = next + (int)threadIdx.x;
        if ((int)threadIdx.x < step && pos < cend) {
            const int ci = g.chunks ? g.chunks[pos] : pos;
            const ChunkBox B = boxes[ci];
            unsigned mask = 0;
            if (SHARED && ARC) {
                // lines from a common origin: the box's distance range once per chunk.
                // The hardware square root (within an ulp or two) is covered by the
                // margin m, a relative 1e-6.
                if (B.hx >= 0.0f) {
                    const float dx = B.cx - org_x, dy = B.cy - org_y, dz = B.cz - org_z;
                    const float hn = __builtin_amdgcn_sqrtf(B.hx * B.hx + B.hy * B.hy + B.hz * B.hz);
                    const float dn = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
                    const float m = 1e-6f * (fabsf(B.cx) + fabsf(B.cy) + fabsf(B.cz) + fabsf(org_x) + fabsf(org_y) +
                                             fabsf(org_z) + dn + hn) + 1e-30f;
                    const float rmin = fmaxf(0.0f, dn - hn) - m, rmax = dn + hn + m;
                    const bool pos_ok = rmax > r_lo && rmin < r_hi;
                    const bool neg_ok = -rmin > r_lo && -rmax < r_hi;
                    CullParams cpc = cp;
                    bool any = pos_ok || neg_ok;
                    if (LUT && any && c.prof.slice_box && cp.lut_dl > 0.0f && cp.lut_de > 0.0f) {
                        // the records the chunk's distances can use: slices of r in
                        // [-rmax, -rmin] and/or [rmin, rmax], one slice of slack each way
                        const ProfileParams& P = c.prof;
                        const float lo = neg_ok ? -rmax : rmin, hi = pos_ok ? rmax : -rmin;
                        const float top = (float)(P.nr - 2);
                        const float t0 = fminf(fmaxf(floorf((lo - P.r_min) * P.inv_dr) - 1.0f, 0.0f), top);
                        const float t1 = fminf(fmaxf(floorf((hi - P.r_min) * P.inv_dr) + 1.0f, 0.0f), top);
                        any = narrow_support(P, cpc, (int)t0, (int)t1);
                    }
                    if (any) {
                        for (int k = 0; k < nl; k++)
                            if (chunk_hits_shared<LUT>(cpc, tl[k], B, dx, dy, dz, m, pos_ok, neg_ok)) mask |= 1u << k;
                    }
                }
            } else if (SPLINE && ss.iboxes) {
                // each line against the box of the chunk's control points on its knot interval
                const ChunkBox* ib = ss.iboxes + (size_t)ci * ss.n_iv;
                for (int k = 0; k < nl; k++)
                    if (chunk_hits<ARC, LUT>(cp, r_lo, r_hi, tl[k], ib[tb[k].lo])) mask |= 1u << k;
            } else {
                for (int k = 0; k < nl; k++)
                    if (chunk_hits<ARC, LUT>(cp, r_lo, r_hi, tl[k], B)) mask |= 1u << k;
            }
            {
                if (mask) {
                    const int slot = atomicAdd(&list_n, 1);
                    if (g.list_packed) {
                        list_word[slot] = ((unsigned)ci << 4) | mask;
                    } else {
                        list_chunk[slot] = ci;
                        list_mask[slot] = mask;
                    }
                    hits += __popc(mask);
                }
            }
        }
        next += step;
        __syncthreads();
        const int n = list_n;
        __syncthreads();
        if (next < cend && n <= g.list_cap - step) continue;
        // (B) project the listed chunks, one wave per chunk
        for (int j = wave; j < n; j += nwaves) {
            // one chunk per wave: its index and hit mask are wave-uniform, so the
            // loop over hit lines is scalar and the line data is a broadcast LDS read
            int ci;
            unsigned mask;
            if (g.list_packed) {
                const unsigned w = __builtin_amdgcn_readfirstlane(list_word[j]);
                ci = (int)(w >> 4);
                mask = w & 15u;
            } else {
                ci = __builtin_amdgcn_readfirstlane(list_chunk[j]);
                mask = __builtin_amdgcn_readfirstlane(list_mask[j]);
            }
            const int s = ci * kChunk + lane;
            if (s >= N) continue;
            if (SPLINE) {
                const float amp = FIXED ? ss.amp[s] * c.amp_scale : ss.amp[s];
                const size_t pitch = (size_t)N;
                while (mask) {
                    const int k = __builtin_ctz(mask);
                    mask &= mask - 1;
                    const SplineLineBasis& Bk = tb[k];
                    // CpuAlgorithm.cpp:137-140: float sum from zero, in control-point order
                    float px = 0.0f, py = 0.0f, pz = 0.0f;
                    if (CP_AHEAD) {
                        // every active control point's loads before the first is used: one
                        // memory latency per hit, not one per control point
                        const int count = __builtin_amdgcn_readfirstlane(Bk.count);  // one line: wave-uniform
                        const size_t base = (size_t)__builtin_amdgcn_readfirstlane(Bk.lo) * pitch + s;
                        float cx[kMaxSplineDegree + 1], cy[kMaxSplineDegree + 1], cz[kMaxSplineDegree + 1];
#pragma unroll
                        for (int i = 0; i <= kMaxSplineDegree; i++) {
                            if (i < count) {
                                const size_t idx = base + (size_t)i * pitch;
                                cx[i] = ss.cx[idx];
                                cy[i] = ss.cy[idx];
                                cz[i] = ss.cz[idx];
                            }
                        }
#pragma unroll
                        for (int i = 0; i <= kMaxSplineDegree; i++) {
                            if (i < count) {
                                px = px + cx[i] * Bk.b[i];
                                py = py + cy[i] * Bk.b[i];
                                pz = pz + cz[i] * Bk.b[i];
                            }
                        }
                    } else {
                        for (int i = 0; i < Bk.count; i++) {
                            const size_t idx = (size_t)(Bk.lo + i) * pitch + s;
                            px = px + ss.cx[idx] * Bk.b[i];
                            py = py + ss.cy[idx] * Bk.b[i];
                            pz = pz + ss.cz[idx] * Bk.b[i];
                        }
                    }
                    const Local q = local_coords<ARC>(tl[k], px, py, pz);
                    double ti;
                    const int bin = radial_bin(c, q.r, ti);
                    emit<PHASE, LUT>(c, bins + k * stride, b0, nb, q.r, q.l, q.e, bin, ti, amp, tally);
                }
            } else if (SHARED && ARC) {
                const float4 p = fs.pts[s];
                const float amp = FIXED ? p.w * c.amp_scale : p.w;
                const float dx = p.x - org_x, dy = p.y - org_y, dz = p.z - org_z;
                const float absd = __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
                double ti_pos;
                const int bin_pos = radial_bin(c, absd, ti_pos);
                const int lb_pos = bin_pos - b0;
                const bool in_pos = bin_pos >= 0 && lb_pos >= 0 && lb_pos < nb;
                LutRadial R{};
                if (LUT) R = lut_radial(c.prof, absd);
                // phase of the positive-side bin: one sincos per scatterer, not per line
                float co_pos = 1.0f, s_pos = 0.0f;
                if (PHASE && in_pos) sincosf((float)((double)bin_pos - ti_pos) * c.phase_per_sample, &s_pos, &co_pos);
                auto hit = [&](int k, float rx, float ry, float rz, float lx, float ly, float lz, float ex, float ey,
                               float ez) {
                    const float rp = dx * rx + dy * ry + dz * rz;
                    const float l = dx * lx + dy * ly + dz * lz;
                    const float e = dx * ex + dy * ey + dz * ez;
                    if (!signbit(rp)) {
                        if (!in_pos) return;
                        const float w =
                            (LUT ? lut_weight_le(c.prof, R, l, e, tally) : gauss_weight(c.prof, l, e, tally)) * amp;
                        add_bin_rot<PHASE>(bins + k * stride, lb_pos, w, co_pos, s_pos, tally);
                    } else {
                        double ti;
                        const int bin = radial_bin(c, -absd, ti);
                        emit<PHASE, LUT>(c, bins + k * stride, b0, nb, -absd, l, e, bin, ti, amp, tally);
                    }
                };
                if (reg_lines) {
#pragma unroll
                    for (int k = 0; k < kRegLines; k++) {
                        if (mask & (1u << k)) {
                            const LineDirs& D = G[k];
                            hit(k, D.rx, D.ry, D.rz, D.lx, D.ly, D.lz, D.ex, D.ey, D.ez);
                        }
                    }
                } else {
                    while (mask) {
                        const int k = __builtin_ctz(mask);
                        mask &= mask - 1;
                        // all three coordinates before branching: one broadcast read of the line
                        const LineGeom& L = tl[k];
                        hit(k, L.rx, L.ry, L.rz, L.lx, L.ly, L.lz, L.ex, L.ey, L.ez);
                    }
                }
            } else {
                const float4 p = fs.pts[s];
                const float amp = FIXED ? p.w * c.amp_scale : p.w;
                while (mask) {
                    const int k = __builtin_ctz(mask);
                    mask &= mask - 1;
                    const Local q = local_coords<ARC>(tl[k], p.x, p.y, p.z);
                    double ti;
                    const int bin = radial_bin(c, q.r, ti);
                    emit<PHASE, LUT>(c, bins + k * stride, b0, nb, q.r, q.l, q.e, bin, ti, amp, tally);
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) list_n = 0;
        __syncthreads();
    }
    if (g.work) {
        // the block's work, for the order of the next launch: one atomic per wave
        for (int off = 32; off > 0; off >>= 1) hits += __shfl_xor(hits, off);
        if (lane == 0 && hits) atomicAdd(&g.work[wg], hits);
    }
    if (g.stats) {
        unsigned long long ev = (unsigned)tally.evals, ad = (unsigned)tally.adds, hc = (unsigned)hits;
        for (int off = 32; off > 0; off >>= 1) {
            ev += __shfl_xor(ev, off);
            ad += __shfl_xor(ad, off);
            if (!g.work) hc += __shfl_xor(hc, off);  // (already summed over the wave for the block order)
        }
        if (lane == 0) {
            atomicAdd(&wg_stats[0], ev);
            atomicAdd(&wg_stats[1], ad);
            atomicAdd(&wg_stats[2], hc);
        }
    }
    __syncthreads();
    if (g.stats && threadIdx.x < 3) atomicAdd(&g.stats[threadIdx.x], wg_stats[threadIdx.x]);  // one per block
    // flush: line k's window [b0, b0 + nb) of the global bins [line][num_rf]
    const int row = nb * E;
    for (int i = threadIdx.x; i < nl * row; i += blockDim.x) {
        const int k = i / row, o = i - k * row;
        const float v = bin_value(bins[k * stride + o], c);
        float* dst = out + ((size_t)(line0 + k) * c.num_rf + b0) * E + o;
        if (g.atomic) {
            if (v != 0.0f) unsafeAtomicAdd(dst, v);
        } else {
            *dst = v;
        }
    }
}

template <bool ARC, bool PHASE, bool LUT, bool SPLINE, bool SHARED, bool FIXED>
__global__ __launch_bounds__(1024) void project_tile_kernel(ProjCommon c, CullParams cp, FixedSet fs, SplineSet ss,
                                                            const ChunkBox* __restrict__ boxes, int nchunks,
                                                            float* __restrict__ out, TileLaunch g) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ int list_n;
    const int wg = g.order ? g.order[blockIdx.x] : (int)blockIdx.x;
    tile_block<ARC, PHASE, LUT, SPLINE, SHARED, FIXED>(c, cp, fs, ss, boxes, nchunks, out, g, wg, blockIdx.y, lds,
                                                       list_n);
}

// Order of the next tile launch's blocks: by the (chunk, line) pairs each listed
// in the previous launch, heaviest first, so that the heaviest blocks start in
// the first wave of workgroups and the launch does not end on a long tail; the
// counts are then cleared for the next launch to fill.  A counting sort into
// kOrderBuckets buckets spaced evenly up to the largest count: a few barriers
// instead of a full sort (order within a bucket is arbitrary).
constexpr int kOrderBuckets = 256;

// place(i, r): block i goes to position r of the order.
template <class Place>
__device__ __forceinline__ void order_blocks(int* __restrict__ work, int n, Place place) {
    __shared__ int count[kOrderBuckets];
    __shared__ int start[kOrderBuckets];
    __shared__ int wmax;
    if (threadIdx.x < kOrderBuckets) count[threadIdx.x] = 0;
    if (threadIdx.x == 0) wmax = 1;
    __syncthreads();
    int m = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = max(m, work[i]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&wmax, m);
    __syncthreads();
    const long long top = wmax;
    // heaviest first: bucket 0 holds the largest counts
    const auto bucket = [top](int w) {
        return (kOrderBuckets - 1) - (int)(((long long)max(w, 0) * (kOrderBuckets - 1)) / top);
    };
    for (int i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&count[bucket(work[i])], 1);
    __syncthreads();
    if (threadIdx.x < 64) {
        // exclusive scan of the counts: four buckets per lane, then across the wave
        constexpr int Q = kOrderBuckets / 64;
        const int l = threadIdx.x;
        int c[Q], sum = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) sum += (c[q] = count[Q * l + q]);
        int incl = sum;
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d);
            if (l >= d) incl += t;
        }
        int run = incl - sum;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            start[Q * l + q] = run;
            run += c[q];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int w = work[i];
        place(i, atomicAdd(&start[bucket(w)], 1));
        work[i] = 0;
    }
}

__global__ __launch_bounds__(1024) void order_tiles_kernel(int* __restrict__ work, int* __restrict__ order, int n) {
    order_blocks(work, n, [order](int i, int r) { order[r] = i; });
}

// A batched job's items are (job, slice * n + block), slice-major; the order is
// written into them (the block of item slice * n + r is the r-th heaviest), so
// the persistent kernel takes the job's heaviest blocks first of every slice.
__global__ __launch_bounds__(1024) void order_segments_kernel(const OrderSeg* __restrict__ segs) {
    const OrderSeg S = segs[blockIdx.x];
    order_blocks(S.work, S.n, [S](int i, int r) {
        for (int s = 0; s < S.slices; s++) S.items[(size_t)s * S.n + r].y = s * S.n + i;
    });
}

hipError_t launch_order_tiles(int* work, int* order, int n, hipStream_t stream) {
    if (n <= 0 || n > kMaxOrderedTiles) return hipErrorInvalidValue;
    hipLaunchKernelGGL(order_tiles_kernel, dim3(1), dim3(1024), 0, stream, work, order, n);
    return hipGetLastError();
}

hipError_t launch_order_segments(const OrderSeg* segs, int n_segs, hipStream_t stream) {
    if (n_segs <= 0) return hipSuccess;
    hipLaunchKernelGGL(order_segments_kernel, dim3(n_segs), dim3(1024), 0, stream, segs);
    return hipGetLastError();
}

// Persistent projection over many jobs (simulators x datasets) of one kernel
// variant: every workgroup takes (job, block) items from a global counter until
// the list is exhausted, and runs the block.  KIND: 0 fixed points, 1 fixed
// points with a shared origin (arc projection), 2 splines; FIXED: fixed-point bins.
template <bool ARC, bool PHASE, bool LUT, int KIND, bool FIXED>
__global__ __launch_bounds__(1024) void project_batch_kernel(const ProjJob* __restrict__ jobs,
                                                             const int2* __restrict__ items, int n_items,
                                                             int* __restrict__ counter) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ int list_n;
    __shared__ int item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(item_s);
        __syncthreads();
        if (it >= n_items) break;
        const int2 w = items[it];  // (job, block within the job)
        const ProjJob& J = jobs[__builtin_amdgcn_readfirstlane(w.x)];
        const int blk = __builtin_amdgcn_readfirstlane(w.y);
        const int per_slice = J.blocks_per_slice;
        const int wg = blk % per_slice, slice = blk / per_slice;
        tile_block<ARC, PHASE, LUT, KIND == 2, KIND == 1, FIXED>(J.c, J.cp, J.fs, J.ss, J.boxes, J.nchunks, J.bins,
                                                                 J.g, wg, slice, lds, list_n);
        __syncthreads();  // the next item reuses the LDS
    }
}

template <bool ARC, bool PHASE, bool LUT, int KIND, bool FIXED>
static hipError_t launch_batch_t(const ProjJob* jobs, const int2* items, int n_items, int* counter, size_t lds,
                                 int workgroups, hipStream_t stream) {
    auto k = project_batch_kernel<ARC, PHASE, LUT, KIND, FIXED>;
    hipError_t e = allow_lds((const void*)k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(workgroups), dim3(1024), lds, stream, jobs, items, n_items, counter);
    return hipGetLastError();
}

template <bool ARC, bool PHASE, bool LUT>
static hipError_t launch_batch_kind(int kind, bool fixed, const ProjJob* jobs, const int2* items, int n_items,
                                    int* counter, size_t lds, int workgroups, hipStream_t stream) {
#define BCSIM_BATCH_K(K)                                                                                          \
    return fixed ? launch_batch_t<ARC, PHASE, LUT, K, true>(jobs, items, n_items, counter, lds, workgroups, stream) \
                 : launch_batch_t<ARC, PHASE, LUT, K, false>(jobs, items, n_items, counter, lds, workgroups, stream)
    if (kind == 2) BCSIM_BATCH_K(2);
    if (kind == 1) BCSIM_BATCH_K(1);
    BCSIM_BATCH_K(0);
#undef BCSIM_BATCH_K
}

hipError_t launch_project_batch(int variant, const ProjJob* jobs, const int2* items, int n_items, int* counter,
                                size_t lds, int workgroups, hipStream_t stream) {
    if (n_items == 0) return hipSuccess;
    const bool arc = variant & kJobArc, phase = variant & kJobPhase, lut = variant & kJobLut;
    const int kind = (variant & kJobSpline) ? 2 : (arc && (variant & kJobShared)) ? 1 : 0;
    const bool fixed = variant & kJobFixed;
#define BCSIM_BATCH(A, P, U) \
    return launch_batch_kind<A, P, U>(kind, fixed, jobs, items, n_items, counter, lds, workgroups, stream)
    if (arc) {
        if (phase) {
            if (lut) BCSIM_BATCH(true, true, true);
            BCSIM_BATCH(true, true, false);
        }
        if (lut) BCSIM_BATCH(true, false, true);
        BCSIM_BATCH(true, false, false);
    }
    if (phase) {
        if (lut) BCSIM_BATCH(false, true, true);
        BCSIM_BATCH(false, true, false);
    }
    if (lut) BCSIM_BATCH(false, false, true);
    BCSIM_BATCH(false, false, false);
#undef BCSIM_BATCH
}

size_t tile_lds_bytes(int window_bins, size_t bin_bytes, int tile_lines, int list_cap, int list_packed) {
    return (size_t)tile_lines * ((size_t)window_bins * bin_bytes + kTileLineLds) +
           (size_t)list_cap * (list_packed ? 4 : 8);
}

template <bool ARC, bool PHASE, bool LUT, bool SPLINE, bool SHARED, bool FIXED>
static hipError_t launch_tile_f(const ProjCommon& c, const CullParams& cp, const FixedSet& fs, const SplineSet& ss,
                                const ChunkBox* boxes, int nchunks, float* bins, const TileLaunch& cfg,
                                hipStream_t stream) {
    const int tiles = (c.num_lines + cfg.tile_lines - 1) / cfg.tile_lines;
    const size_t lds =
        tile_lds_bytes(cfg.window_bins, tile_bin_bytes(PHASE, FIXED), cfg.tile_lines, cfg.list_cap, cfg.list_packed);
    auto k = project_tile_kernel<ARC, PHASE, LUT, SPLINE, SHARED, FIXED>;
    hipError_t e = allow_lds((const void*)k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(tiles * cfg.windows, cfg.slices), dim3(cfg.block), lds, stream, c, cp, fs, ss, boxes,
                       nchunks, bins, cfg);
    return hipGetLastError();
}

template <bool ARC, bool PHASE, bool LUT, bool SPLINE, bool SHARED>
static hipError_t launch_tile_t(const ProjCommon& c, const CullParams& cp, const FixedSet& fs, const SplineSet& ss,
                                const ChunkBox* boxes, int nchunks, float* bins, const TileLaunch& cfg,
                                hipStream_t stream) {
    if (cfg.fixed)
        return launch_tile_f<ARC, PHASE, LUT, SPLINE, SHARED, true>(c, cp, fs, ss, boxes, nchunks, bins, cfg, stream);
    return launch_tile_f<ARC, PHASE, LUT, SPLINE, SHARED, false>(c, cp, fs, ss, boxes, nchunks, bins, cfg, stream);
}

template <bool ARC, bool PHASE, bool LUT>
static hipError_t launch_tile_kind(const ProjCommon& c, const CullParams& cp, const FixedSet* fixed,
                                   const SplineSet* spline, const ChunkBox* boxes, int nchunks, float* bins,
                                   bool shared, const TileLaunch& cfg, hipStream_t stream) {
    const FixedSet fs = fixed ? *fixed : FixedSet{nullptr, 0};
    const SplineSet ss = spline ? *spline : SplineSet{nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, 0};
    if (spline) return launch_tile_t<ARC, PHASE, LUT, true, false>(c, cp, fs, ss, boxes, nchunks, bins, cfg, stream);
    if (ARC && shared)
        return launch_tile_t<ARC, PHASE, LUT, false, true>(c, cp, fs, ss, boxes, nchunks, bins, cfg, stream);
    return launch_tile_t<ARC, PHASE, LUT, false, false>(c, cp, fs, ss, boxes, nchunks, bins, cfg, stream);
}

hipError_t launch_project_tile(const ProjCommon& c, const CullParams& cp, const FixedSet* fixed,
                               const SplineSet* spline, const ChunkBox* boxes, int nchunks, void* bins_v,
                               bool arc, bool phase, bool lut, bool shared_origin, const TileLaunch& cfg,
                               hipStream_t stream) {
    if (nchunks == 0) return hipSuccess;
    float* bins = static_cast<float*>(bins_v);
#define BCSIM_TILE(A, P, U) \
    return launch_tile_kind<A, P, U>(c, cp, fixed, spline, boxes, nchunks, bins, shared_origin, cfg, stream)
    if (arc) {
        if (phase) {
            if (lut) BCSIM_TILE(true, true, true);
            BCSIM_TILE(true, true, false);
        }
        if (lut) BCSIM_TILE(true, false, true);
        BCSIM_TILE(true, false, false);
    }
    if (phase) {
        if (lut) BCSIM_TILE(false, true, true);
        BCSIM_TILE(false, true, false);
    }
    if (lut) BCSIM_TILE(false, false, true);
    BCSIM_TILE(false, false, false);
#undef BCSIM_TILE
}

}  // namespace bcsim_hip
