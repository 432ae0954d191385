// Culled tile projection for gfx950: the production scatterer-projection kernel.
//
// Work decomposition.  grid.x walks tiles of `tile_lines` scan lines whose bins
// live in LDS for the whole kernel (up to ~156 KiB of the 160 KiB); grid.y
// splits the scatterer chunks into slices.  The host stores every dataset in
// Morton order of position and keeps one axis-aligned box per chunk of 64
// scatterers (ChunkBox), so a chunk is spatially compact.
//
// Per batch of up to kListCap chunks a workgroup
//   (A) tests every chunk's box against the beam support of each line of its
//       tile -- the region outside which the weight is exactly zero in float
//       (LUT table extent, Gaussian underflow, bin window) -- and appends the
//       surviving chunks with a per-line hit mask to a list in LDS;
//   (B) hands the listed chunks to its waves, one chunk per wave, one scatterer
//       per lane; each lane projects its scatterer onto every hit line of the
//       tile and adds into the LDS bins with ds_add_f32.
// The tile leaves LDS once: plain stores when the workgroup saw every chunk,
// otherwise atomic adds of the non-zero bins.
//
// Skipped pairs would have added w*a == 0 to a bin, so the sums are those of the
// direct evaluation up to float summation order.  When every line starts at the
// same origin (sector and pyramidal scans) the arc-projected distance |p - o|
// and its bin are computed once per scatterer instead of once per line.
//
// Compiled with -ffp-contract=off (see proj_math.h).
#include <hip/hip_runtime.h>

#include "proj_math.h"
#include "sim_device.h"

namespace bcsim_hip {

// Can any point of box B have a non-zero weight and a valid bin on line L?
// Conservative: bounds are padded by a relative margin that covers the float
// rounding of the per-point evaluation.
template <bool ARC, bool LUT>
__device__ __forceinline__ bool chunk_hits(const CullParams& cp, const LineGeom& L, const ChunkBox& B) {
    if (!(B.hx >= 0.0f)) return false;  // empty chunk
    const float dx = B.cx - L.ox, dy = B.cy - L.oy, dz = B.cz - L.oz;
    const float hn = sqrtf(B.hx * B.hx + B.hy * B.hy + B.hz * B.hz);
    const float dn = sqrtf(dx * dx + dy * dy + dz * dz);
    const float m = 1e-6f * (fabsf(B.cx) + fabsf(B.cy) + fabsf(B.cz) + fabsf(L.ox) + fabsf(L.oy) + fabsf(L.oz) +
                             dn + hn) + 1e-30f;
    const float rc = dx * L.rx + dy * L.ry + dz * L.rz;
    const float hr = B.hx * fabsf(L.rx) + B.hy * fabsf(L.ry) + B.hz * fabsf(L.rz) + m;
    const float lc = dx * L.lx + dy * L.ly + dz * L.lz;
    const float hl = B.hx * fabsf(L.lx) + B.hy * fabsf(L.ly) + B.hz * fabsf(L.lz) + m;
    const float ec = dx * L.ex + dy * L.ey + dz * L.ez;
    const float he = B.hx * fabsf(L.ex) + B.hy * fabsf(L.ey) + B.hz * fabsf(L.ez) + m;
    bool hit_r;
    if (ARC) {
        // r = +-|p - o| with the sign of the projection onto the beam direction
        const float rmin = fmaxf(0.0f, dn - hn) - m, rmax = dn + hn + m;
        const bool pos = (rc + hr > 0.0f) && (rmax > cp.r_lo) && (rmin < cp.r_hi);
        const bool neg = (rc - hr < 0.0f) && (-rmin > cp.r_lo) && (-rmax < cp.r_hi);
        hit_r = pos || neg;
    } else {
        hit_r = (rc + hr > cp.r_lo) && (rc - hr < cp.r_hi);
    }
    if (!hit_r) return false;
    if (LUT) return (lc + hl > cp.l_lo) && (lc - hl < cp.l_hi) && (ec + he > cp.e_lo) && (ec - he < cp.e_hi);
    const float dl = fmaxf(0.0f, fabsf(lc) - hl), de = fmaxf(0.0f, fabsf(ec) - he);
    return !(cp.g_a * dl * dl + cp.g_b * de * de >= cp.g_max);
}

template <bool ARC, bool PHASE, bool LUT, bool SPLINE, bool SHARED>
__global__ __launch_bounds__(1024) void project_tile_kernel(ProjCommon c, CullParams cp, FixedSet fs, SplineSet ss,
                                                            const ChunkBox* __restrict__ boxes, int nchunks,
                                                            float* __restrict__ out, int tile_lines,
                                                            int chunks_per_slice, int atomic) {
    extern __shared__ float lds[];
    __shared__ int list_n;
    constexpr int E = PHASE ? 2 : 1;
    const int line0 = blockIdx.x * tile_lines;
    const int nl = min(tile_lines, c.num_lines - line0);
    const int per_line = c.num_rf * E;
    float* bins = lds;
    int* list_chunk = reinterpret_cast<int*>(lds + (size_t)tile_lines * per_line);
    unsigned* list_mask = reinterpret_cast<unsigned*>(list_chunk + kListCap);
    zero_lds(bins, nl * per_line);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    const int cbeg = blockIdx.y * chunks_per_slice;
    const int cend = min(nchunks, cbeg + chunks_per_slice);
    const int N = SPLINE ? ss.n : fs.n;

    for (int b0 = cbeg; b0 < cend; b0 += kListCap) {
        if (threadIdx.x == 0) list_n = 0;
        __syncthreads();
        // (A) cull
        const int b1 = min(cend, b0 + kListCap);
        for (int ci = b0 + (int)threadIdx.x; ci < b1; ci += blockDim.x) {
            const ChunkBox B = boxes[ci];
            unsigned mask = 0;
            for (int k = 0; k < nl; k++)
                if (chunk_hits<ARC, LUT>(cp, c.lines[line0 + k], B)) mask |= 1u << k;
            if (mask) {
                const int slot = atomicAdd(&list_n, 1);
                list_chunk[slot] = ci;
                list_mask[slot] = mask;
            }
        }
        __syncthreads();
        // (B) project the surviving chunks, one wave per chunk
        const int n = list_n;
        for (int j = wave; j < n; j += nwaves) {
            const int ci = list_chunk[j];
            const unsigned mask = list_mask[j];
            const int s = ci * kChunk + lane;
            if (s >= N) continue;
            if (SPLINE) {
                const float amp = ss.amp[s];
                const size_t stride = (size_t)N;
                for (int k = 0; k < nl; k++) {
                    if (!((mask >> k) & 1u)) continue;
                    const SplineLineBasis& Bk = ss.basis[line0 + k];
                    // CpuAlgorithm.cpp:137-140: float sum from zero, in control-point order
                    float px = 0.0f, py = 0.0f, pz = 0.0f;
                    for (int i = 0; i < Bk.count; i++) {
                        const size_t idx = (size_t)(Bk.lo + i) * stride + s;
                        px = px + ss.cx[idx] * Bk.b[i];
                        py = py + ss.cy[idx] * Bk.b[i];
                        pz = pz + ss.cz[idx] * Bk.b[i];
                    }
                    project_point<ARC, PHASE, LUT>(c, c.lines[line0 + k], px, py, pz, amp, bins + k * per_line);
                }
            } else if (SHARED && ARC) {
                const float4 p = fs.pts[s];
                const LineGeom& L0 = c.lines[line0];
                const float dx = p.x - L0.ox, dy = p.y - L0.oy, dz = p.z - L0.oz;
                const float absd = __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
                double ti_pos;
                const int bin_pos = radial_bin(c, absd, ti_pos);
                for (int k = 0; k < nl; k++) {
                    if (!((mask >> k) & 1u)) continue;
                    const LineGeom& L = c.lines[line0 + k];
                    const float rp = dx * L.rx + dy * L.ry + dz * L.rz;
                    const float l = dx * L.lx + dy * L.ly + dz * L.lz;
                    const float e = dx * L.ex + dy * L.ey + dz * L.ez;
                    if (!signbit(rp)) {
                        accumulate<PHASE, LUT>(c, absd, l, e, bin_pos, ti_pos, p.w, bins + k * per_line);
                    } else {
                        double ti;
                        const int bin = radial_bin(c, -absd, ti);
                        accumulate<PHASE, LUT>(c, -absd, l, e, bin, ti, p.w, bins + k * per_line);
                    }
                }
            } else {
                const float4 p = fs.pts[s];
                for (int k = 0; k < nl; k++) {
                    if (!((mask >> k) & 1u)) continue;
                    project_point<ARC, PHASE, LUT>(c, c.lines[line0 + k], p.x, p.y, p.z, p.w, bins + k * per_line);
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    flush_bins(bins, out + (size_t)line0 * per_line, nl * per_line, atomic != 0);
}

size_t tile_lds_bytes(int num_rf, bool complex_bins, int tile_lines) {
    return (size_t)tile_lines * (size_t)num_rf * (complex_bins ? 8 : 4) + (size_t)kListCap * 8;
}

template <bool ARC, bool PHASE, bool LUT, bool SPLINE, bool SHARED>
static hipError_t launch_tile_t(const ProjCommon& c, const CullParams& cp, const FixedSet& fs, const SplineSet& ss,
                                const ChunkBox* boxes, int nchunks, float* bins, const TileLaunch& cfg,
                                hipStream_t stream) {
    const int tiles = (c.num_lines + cfg.tile_lines - 1) / cfg.tile_lines;
    const size_t lds = tile_lds_bytes(c.num_rf, PHASE, cfg.tile_lines);
    auto k = project_tile_kernel<ARC, PHASE, LUT, SPLINE, SHARED>;
    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(tiles, cfg.slices), dim3(cfg.block), lds, stream, c, cp, fs, ss, boxes, nchunks,
                       bins, cfg.tile_lines, cfg.chunks_per_slice, cfg.atomic);
    return hipGetLastError();
}

template <bool ARC, bool PHASE, bool LUT>
static hipError_t launch_tile_kind(const ProjCommon& c, const CullParams& cp, const FixedSet* fixed,
                                   const SplineSet* spline, const ChunkBox* boxes, int nchunks, float* bins,
                                   bool shared, const TileLaunch& cfg, hipStream_t stream) {
    const FixedSet fs = fixed ? *fixed : FixedSet{nullptr, 0};
    const SplineSet ss = spline ? *spline : SplineSet{nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0};
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
