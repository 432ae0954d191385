// HIP kernels of the simulation path, written for gfx950 (CDNA4, wave64, 160 KiB LDS).
//
//   project_*_kernel     scatterer projection onto scan lines (reference kernels
//                        FixedAlgKernel, cuda_kernels_fixed.cuh:15-65, and
//                        SplineAlgKernel, cuda_kernels_spline2.cu:14-79; arithmetic
//                        of the CPU algorithm, CpuAlgorithm.cpp:50-183).  A workgroup
//                        holds the bins of a tile of lines in LDS, every scatterer it
//                        reads is projected onto all lines of the tile, and the tile
//                        leaves LDS once, by plain stores or by global atomics when
//                        the scatterers are split over several workgroups.
//   render_spline_kernel evaluate splines at one time (RenderSplineKernel,
//                        cuda_kernels_spline1.cu:15-45).
//   convolve_demod_kernel one line per workgroup, entirely in LDS: forward FFT,
//                        multiply by the Hilbert-masked excitation spectrum, inverse
//                        FFT, delay extraction, demodulation and decimation
//                        (cuFFT + MultiplyFftKernel + cuFFT + DemodulateKernel +
//                        host extraction in GpuAlgorithm.cpp:280-376; CPU semantics
//                        BeamConvolver.cpp:64-73, CpuAlgorithm.cpp:349-365).
//
// This file is compiled with -ffp-contract=off: the radial coordinate and the
// spline position must round exactly like the reference's float code because
// they select a bin.  The FFT code opts back into contraction locally.
#include <hip/hip_runtime.h>

#include <mutex>
#include <set>
#include <utility>

#include "proj_math.h"
#include "sim_device.h"

namespace bcsim_hip {

// ------------------------------------------------------------------------------------
// Scratch of the allocation-free frame entry point: one frame's bins and rendered
// spline points live here instead of in buffers of the simulator.  Every GPU the
// library runs on holds its own copy until the process ends; host code never
// passes the array's name to a kernel (in host code it is only a stand-in), but
// the device address looked up below for the current device.

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(int device, void** ptr) {
    static std::mutex mu;
    static void* addr[kMaxScratchDevices] = {};
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

// ------------------------------------------------------------------------------------
// utilities

hipError_t allow_lds(const void* func, size_t bytes) {
    constexpr size_t kMaxDynamic = (size_t)kLdsBytes - 256;  // below the static LDS of every kernel here
    if (bytes > kMaxDynamic) return hipErrorInvalidValue;
    static std::mutex mu;
    static std::set<std::pair<int, const void*>> raised;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(mu);
    if (raised.count({dev, func})) return hipSuccess;
    e = hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxDynamic);
    if (e == hipSuccess) raised.insert({dev, func});
    return e;
}

__global__ void fill_zero_kernel(float4* p4, size_t n4, float* tail, int n_tail) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        p4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (blockIdx.x == 0 && (int)threadIdx.x < n_tail) tail[threadIdx.x] = 0.f;
}

hipError_t launch_fill_zero(float* p, size_t n_floats, hipStream_t stream) {
    if (n_floats == 0) return hipSuccess;
    const size_t n4 = n_floats / 4;
    const int tail = (int)(n_floats - n4 * 4);
    size_t blocks = (n4 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                       reinterpret_cast<float4*>(p), n4, p + n4 * 4, tail);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// direct projection: every (line, scatterer) pair, scatterers in the caller's
// order (projection_mode "direct"; the culled tile kernel is project_tile.hip)

template <bool ARC, bool PHASE, bool LUT>
__global__ __launch_bounds__(256) void project_fixed_kernel(ProjCommon c, FixedSet s, float* bins,
                                                            int tile_lines, int chunk, int atomic) {
    extern __shared__ float lds[];
    constexpr int E = PHASE ? 2 : 1;
    const int line0 = blockIdx.x * tile_lines;
    const int nl = min(tile_lines, c.num_lines - line0);
    const int per_line = c.num_rf * E;
    zero_lds(lds, nl * per_line);
    __syncthreads();
    const int begin = blockIdx.y * chunk;
    const int end = min(s.n, begin + chunk);
    for (int n = begin + threadIdx.x; n < end; n += blockDim.x) {
        const float4 p = s.pts[n];
        for (int k = 0; k < nl; k++)
            project_point<ARC, PHASE, LUT>(c, c.lines[line0 + k], p.x, p.y, p.z, p.w, lds + k * per_line);
    }
    __syncthreads();
    flush_bins(lds, bins + (size_t)line0 * per_line, nl * per_line, atomic != 0);
}

template <bool ARC, bool PHASE, bool LUT>
__global__ __launch_bounds__(256) void project_spline_kernel(ProjCommon c, SplineSet s, float* bins,
                                                             int tile_lines, int chunk, int atomic) {
    extern __shared__ float lds[];
    constexpr int E = PHASE ? 2 : 1;
    const int line0 = blockIdx.x * tile_lines;
    const int nl = min(tile_lines, c.num_lines - line0);
    const int per_line = c.num_rf * E;
    zero_lds(lds, nl * per_line);
    __syncthreads();
    const int begin = blockIdx.y * chunk;
    const int end = min(s.n, begin + chunk);
    const size_t N = (size_t)s.n;
    for (int n = begin + threadIdx.x; n < end; n += blockDim.x) {
        const float amp = s.amp[n];
        for (int k = 0; k < nl; k++) {
            const SplineLineBasis& B = s.basis[line0 + k];
            // CpuAlgorithm.cpp:137-140: float sum from zero, in control-point order
            float px = 0.0f, py = 0.0f, pz = 0.0f;
            for (int i = 0; i < B.count; i++) {
                const size_t idx = (size_t)(B.lo + i) * N + n;
                px = px + s.cx[idx] * B.b[i];
                py = py + s.cy[idx] * B.b[i];
                pz = pz + s.cz[idx] * B.b[i];
            }
            project_point<ARC, PHASE, LUT>(c, c.lines[line0 + k], px, py, pz, amp, lds + k * per_line);
        }
    }
    __syncthreads();
    flush_bins(lds, bins + (size_t)line0 * per_line, nl * per_line, atomic != 0);
}

#define BCSIM_SELECT(kern, arc, phase, lut)                                                 \
    ((arc) ? ((phase) ? ((lut) ? kern<true, true, true> : kern<true, true, false>)          \
                      : ((lut) ? kern<true, false, true> : kern<true, false, false>))       \
           : ((phase) ? ((lut) ? kern<false, true, true> : kern<false, true, false>)        \
                      : ((lut) ? kern<false, false, true> : kern<false, false, false>)))

static size_t proj_lds_bytes(const ProjCommon& c, bool complex_bins, int tile) {
    return (size_t)tile * (size_t)c.num_rf * (complex_bins ? 8 : 4);
}

hipError_t launch_project_fixed(const ProjCommon& c, const FixedSet& s, void* bins, bool complex_bins,
                                bool arc, bool phase, bool lut, const ProjLaunch& cfg,
                                hipStream_t stream) {
    if (s.n == 0) return hipSuccess;
    const int tiles = (c.num_lines + cfg.tile_lines - 1) / cfg.tile_lines;
    const int chunk = (s.n + cfg.slices - 1) / cfg.slices;
    const size_t lds = proj_lds_bytes(c, complex_bins, cfg.tile_lines);
    auto k = BCSIM_SELECT(project_fixed_kernel, arc, phase, lut);
    hipError_t e = allow_lds((const void*)k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(tiles, cfg.slices), dim3(cfg.block), lds, stream, c, s,
                       static_cast<float*>(bins), cfg.tile_lines, chunk, cfg.atomic);
    return hipGetLastError();
}

hipError_t launch_project_spline(const ProjCommon& c, const SplineSet& s, void* bins, bool complex_bins,
                                 bool arc, bool phase, bool lut, const ProjLaunch& cfg,
                                 hipStream_t stream) {
    if (s.n == 0) return hipSuccess;
    const int tiles = (c.num_lines + cfg.tile_lines - 1) / cfg.tile_lines;
    const int chunk = (s.n + cfg.slices - 1) / cfg.slices;
    const size_t lds = proj_lds_bytes(c, complex_bins, cfg.tile_lines);
    auto k = BCSIM_SELECT(project_spline_kernel, arc, phase, lut);
    hipError_t e = allow_lds((const void*)k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(tiles, cfg.slices), dim3(cfg.block), lds, stream, c, s,
                       static_cast<float*>(bins), cfg.tile_lines, chunk, cfg.atomic);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// spline rendering at one time (all lines share a timestamp)

__global__ __launch_bounds__(256) void render_spline_kernel(SplineSet s, SplineLineBasis B, float4* out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= s.n) return;
    const size_t N = (size_t)s.n;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    for (int i = 0; i < B.count; i++) {
        const size_t idx = (size_t)(B.lo + i) * N + n;
        px = px + s.cx[idx] * B.b[i];
        py = py + s.cy[idx] * B.b[i];
        pz = pz + s.cz[idx] * B.b[i];
    }
    out[n] = make_float4(px, py, pz, s.amp[n]);
}

hipError_t launch_render_spline(const SplineSet& s, const SplineLineBasis& basis, float4* out,
                                hipStream_t stream) {
    if (s.n == 0) return hipSuccess;
    hipLaunchKernelGGL(render_spline_kernel, dim3((s.n + 255) / 256), dim3(256), 0, stream, s, basis, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// fused convolution + demodulation, one line per workgroup

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
#pragma clang fp contract(fast)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

struct Twiddle {
    const float2* coarse;
    const float2* fine;
    int bits;
    int mask;
    int shift;  // the table holds exp(-2 pi i k / N); a transform of N >> shift points reads every 2^shift-th
    // exp(-2 pi i idx / (N >> shift)) for 0 <= idx < N >> shift, conjugated for the inverse transform
    template <bool INV>
    __device__ __forceinline__ float2 get(int idx) const {
        idx <<= shift;
        const float2 w = cmul(coarse[idx >> bits], fine[idx & mask]);
        return INV ? make_float2(w.x, -w.y) : w;
    }
};

// 4-point DFT (forward: exp(-2 pi i nk / 4); inverse: conjugate kernel).
template <bool INV>
__device__ __forceinline__ void dft4(float2 u0, float2 u1, float2 u2, float2 u3, float2& y0, float2& y1,
                                     float2& y2, float2& y3) {
    const float2 a0 = cadd(u0, u2);
    const float2 a1 = csub(u0, u2);
    const float2 a2 = cadd(u1, u3);
    const float2 a3 = rot90<INV>(csub(u1, u3));
    y0 = cadd(a0, a2);
    y2 = csub(a0, a2);
    y1 = cadd(a1, a3);
    y3 = csub(a1, a3);
}

// 8-point DFT as two 4-point DFTs of the even and odd inputs joined by the
// eighth roots of unity: X[k] = E[k] + w^k O[k], X[k+4] = E[k] - w^k O[k].
template <bool INV>
__device__ __forceinline__ void dft8(const float2* x, float2* y) {
    constexpr float h = 0.70710678118654752f;  // 1/sqrt(2)
    float2 e[4], o[4];
    dft4<INV>(x[0], x[2], x[4], x[6], e[0], e[1], e[2], e[3]);
    dft4<INV>(x[1], x[3], x[5], x[7], o[0], o[1], o[2], o[3]);
    // w = exp(-i pi/4) forward, exp(+i pi/4) inverse
    const float2 o1 = INV ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                          : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    const float2 o2 = rot90<INV>(o[2]);
    const float2 o3 = INV ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                          : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    y[0] = cadd(e[0], o[0]);
    y[4] = csub(e[0], o[0]);
    y[1] = cadd(e[1], o1);
    y[5] = csub(e[1], o1);
    y[2] = cadd(e[2], o2);
    y[6] = csub(e[2], o2);
    y[3] = cadd(e[3], o3);
    y[7] = csub(e[3], o3);
}

// One Stockham autosort pass of radix R over N points: input i is src(i), output
// i goes to dst(i, value).  Ls is the length of the sub-transforms already
// formed.  With SYNC every thread reads all its inputs before the workgroup
// synchronises and any thread writes (a pass in place over one LDS buffer).
// CN, CLS: N and Ls as compile-time constants (0: the run-time arguments), which
// turns the index arithmetic of the pass into constants and LDS offset fields.
template <int R, int NT, int NMAX, bool INV, bool SYNC, int CN = 0, int CLS = 0, class Src, class Dst>
__device__ __forceinline__ void stockham_io(int N_rt, int Ls_rt, const Twiddle& tw, const Src& src,
                                            const Dst& dst) {
#pragma clang fp contract(fast)
    const int N = CN ? CN : N_rt;
    const int Ls = CLS ? CLS : Ls_rt;
    constexpr int IT = ((CN ? CN : NMAX) / R + NT - 1) / NT;
    const int m = N / R;
    float2 x[IT][R];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int j = threadIdx.x + it * NT;
        if (j < m) {
#pragma unroll
            for (int q = 0; q < R; q++) x[it][q] = src(j + q * m);
        }
    }
    if (SYNC) __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int j = threadIdx.x + it * NT;
        if (j < m) {
            const int k = j & (Ls - 1);
            const int step = k * (N / (Ls * R));
            if (R == 8) {
                // w^1 from the table (two LDS reads), the other powers as products:
                // w^2 = w^1 w^1, w^4 = w^2 w^2 (each product adds about an ulp), then
                // w^3, w^5, w^6, w^7 from those: 2 table reads per butterfly instead of 14
                // (measured: the volume's convolution 0.217 -> 0.208 ms)
                if (Ls > 1) {
                    const float2 w1 = tw.template get<INV>(step);
                    const float2 w2 = cmul(w1, w1);
                    const float2 w4 = cmul(w2, w2);
                    const float2 w3 = cmul(w1, w2);
                    x[it][1] = cmul(x[it][1], w1);
                    x[it][2] = cmul(x[it][2], w2);
                    x[it][3] = cmul(x[it][3], w3);
                    x[it][4] = cmul(x[it][4], w4);
                    x[it][5] = cmul(x[it][5], cmul(w1, w4));
                    x[it][6] = cmul(x[it][6], cmul(w2, w4));
                    x[it][7] = cmul(x[it][7], cmul(w3, w4));
                }
            } else if (R == 4 && Ls > 1) {  // Ls == 1: every twiddle is 1
                const float2 w1 = tw.template get<INV>(step);
                const float2 w2 = cmul(w1, w1);
                x[it][1] = cmul(x[it][1], w1);
                x[it][2] = cmul(x[it][2], w2);
                x[it][3] = cmul(x[it][3], cmul(w1, w2));
            } else if (R == 2 && Ls > 1) {
                x[it][1] = cmul(x[it][1], tw.template get<INV>(step));
            }
            float2 y[R];
            if (R == 2) {
                y[0] = cadd(x[it][0], x[it][1]);
                y[1] = csub(x[it][0], x[it][1]);
            } else if (R == 4) {
                dft4<INV>(x[it][0], x[it][1], x[it][2], x[it][3], y[0], y[1], y[2], y[3]);
            } else {
                dft8<INV>(x[it], y);
            }
            const int base = (j - k) * R + k;
#pragma unroll
            for (int s2 = 0; s2 < R; s2++) dst(base + s2 * Ls, y[s2]);
        }
    }
}

// Radix of the first pass: it takes up log2 n mod 3 (RMAX 8) or log2 n mod 2
// (RMAX 4); every later pass has radix RMAX.
template <int RMAX>
__device__ __forceinline__ int first_radix(int log2n) {
    if (RMAX == 8) {
        const int r = log2n % 3;
        return r == 1 ? 2 : r == 2 ? 4 : 8;
    }
    return (log2n & 1) ? 2 : 4;
}

template <int NT, int NMAX, bool INV, bool SYNC, class Src, class Dst>
__device__ __forceinline__ void stockham_any(int r, int N, int Ls, const Twiddle& tw, const Src& src,
                                             const Dst& dst) {
    if (r == 2) stockham_io<2, NT, NMAX, INV, SYNC>(N, Ls, tw, src, dst);
    else if (r == 4) stockham_io<4, NT, NMAX, INV, SYNC>(N, Ls, tw, src, dst);
    else stockham_io<8, NT, NMAX, INV, SYNC>(N, Ls, tw, src, dst);
}

// Power-of-two FFT of n points through the LDS buffer `buf`: the first pass
// reads its inputs from src0 (SYNC0: src0 reads buf, so the pass is in place),
// the last pass writes its outputs to dstL (SYNCL: dstL writes buf), the passes
// between work in place in buf.  Returns after a workgroup barrier when the
// output is in buf (SYNCL); otherwise without one.
template <int NT, int NMAX, bool INV, int RMAX, bool SYNC0, bool SYNCL, class Src, class Dst>
__device__ __forceinline__ void fft_fused(float2* buf, int n, int log2n, const Twiddle& tw, const Src& src0,
                                          const Dst& dstL) {
    const auto lds_src = [buf](int i) { return buf[conv_slot(i)]; };
    const auto lds_dst = [buf](int i, float2 v) { buf[conv_slot(i)] = v; };
    const int r0 = first_radix<RMAX>(log2n);
    if (n == r0) {  // a single pass
        stockham_any<NT, NMAX, INV, SYNC0 || SYNCL>(r0, n, 1, tw, src0, dstL);
        if (SYNCL) __syncthreads();
        return;
    }
    stockham_any<NT, NMAX, INV, SYNC0>(r0, n, 1, tw, src0, lds_dst);
    __syncthreads();
    int Ls = r0;
    for (; Ls * RMAX < n; Ls *= RMAX) {
        stockham_io<RMAX, NT, NMAX, INV, true>(n, Ls, tw, lds_src, lds_dst);
        __syncthreads();
    }
    stockham_io<RMAX, NT, NMAX, INV, SYNCL>(n, Ls, tw, lds_src, dstL);
    if (SYNCL) __syncthreads();
}

// fft_fused for a transform length 2^LOG2N fixed at compile time: the same
// passes, each with its N and Ls constants.
template <int RMAX>
constexpr int first_radix_c(int log2n) {
    return RMAX == 8 ? (log2n % 3 == 1 ? 2 : log2n % 3 == 2 ? 4 : 8) : ((log2n & 1) ? 2 : 4);
}

template <int NT, int NMAX, bool INV, int RMAX, bool SYNCL, int N, int LS, class Dst>
__device__ __forceinline__ void fft_passes_c(float2* buf, const Twiddle& tw, const Dst& dstL) {
    const auto lds_src = [buf](int i) { return buf[conv_slot(i)]; };
    if constexpr (LS * RMAX < N) {
        const auto lds_dst = [buf](int i, float2 v) { buf[conv_slot(i)] = v; };
        stockham_io<RMAX, NT, NMAX, INV, true, N, LS>(N, LS, tw, lds_src, lds_dst);
        __syncthreads();
        fft_passes_c<NT, NMAX, INV, RMAX, SYNCL, N, LS * RMAX>(buf, tw, dstL);
    } else {
        stockham_io<RMAX, NT, NMAX, INV, SYNCL, N, LS>(N, LS, tw, lds_src, dstL);
        if (SYNCL) __syncthreads();
    }
}

template <int NT, int NMAX, bool INV, int RMAX, bool SYNC0, bool SYNCL, int LOG2N, class Src, class Dst>
__device__ __forceinline__ void fft_fused_c(float2* buf, const Twiddle& tw, const Src& src0, const Dst& dstL) {
    constexpr int n = 1 << LOG2N;
    constexpr int r0 = first_radix_c<RMAX>(LOG2N);
    if constexpr (n == r0) {  // a single pass
        stockham_io<r0, NT, NMAX, INV, SYNC0 || SYNCL, n, 1>(n, 1, tw, src0, dstL);
        if (SYNCL) __syncthreads();
    } else {
        const auto lds_dst = [buf](int i, float2 v) { buf[conv_slot(i)] = v; };
        stockham_io<r0, NT, NMAX, INV, SYNC0, n, 1>(n, 1, tw, src0, lds_dst);
        __syncthreads();
        fft_passes_c<NT, NMAX, INV, RMAX, SYNCL, n, r0>(buf, tw, dstL);
    }
}

// The transform of n = 2^log2n points: the lengths of the volume, sector and
// spline workloads (2^11 .. 2^13, and with W14 2^14) with constant passes, others
// through fft_fused.  W14 only in a kernel for complex transforms of 2^14 points,
// which runs one workgroup per CU anyway (its buffer takes most of the LDS): the
// 2^14-point passes hold more points per thread, and elsewhere their registers
// would cost the second workgroup per CU (measured: the volume's real-bin
// convolution 0.165 -> 0.222 ms with them in its kernel).
template <int NT, int NMAX, bool INV, int RMAX, bool SYNC0, bool SYNCL, bool W14 = false, class Src, class Dst>
__device__ __forceinline__ void fft_any(float2* buf, int n, int log2n, const Twiddle& tw, const Src& src0,
                                        const Dst& dstL) {
    if (W14 && NMAX >= 16384 && log2n == 14)
        fft_fused_c<NT, NMAX, INV, RMAX, SYNC0, SYNCL, 14>(buf, tw, src0, dstL);
    else if (NMAX >= 8192 && log2n == 13)
        fft_fused_c<NT, NMAX, INV, RMAX, SYNC0, SYNCL, 13>(buf, tw, src0, dstL);
    else if (NMAX >= 4096 && log2n == 12)
        fft_fused_c<NT, NMAX, INV, RMAX, SYNC0, SYNCL, 12>(buf, tw, src0, dstL);
    else if (log2n == 11)
        fft_fused_c<NT, NMAX, INV, RMAX, SYNC0, SYNCL, 11>(buf, tw, src0, dstL);
    else
        fft_fused<NT, NMAX, INV, RMAX, SYNC0, SYNCL>(buf, n, log2n, tw, src0, dstL);
}

// Counter-based normal variates for the noise path: a 32-bit integer hash
// (Wang's mix, then a multiply-xorshift round) feeds a Box-Muller transform.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x = (x ^ 61u) ^ (x >> 16);
    x *= 9u;
    x = x ^ (x >> 4);
    x *= 0x27d4eb2du;
    x = x ^ (x >> 15);
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    return x;
}

__device__ __forceinline__ float2 normal_pair(uint32_t seed, uint32_t counter, uint32_t line, uint32_t n) {
    const uint32_t h0 = mix32(seed ^ mix32(counter ^ mix32(line * 0x9e3779b9u ^ mix32(n))));
    const uint32_t h1 = mix32(h0 ^ 0x68bc21ebu);
    const float u1 = ((h0 >> 8) + 1u) * (1.0f / 16777217.0f);  // (0, 1]
    const float u2 = (h1 >> 8) * (1.0f / 16777216.0f);         // [0, 1)
    const float rad = sqrtf(-2.0f * logf(u1));
    float s, co;
    sincospif(2.0f * u2, &s, &co);
    return make_float2(rad * co, rad * s);
}

// One line: bins -> spectrum -> x H -> inverse transform -> demodulate.
//
// Real bins (no noise): the N real samples b[n] are transformed as the N/2
// complex points z[m] = b[2m] + i b[2m+1]; with Z = FFT_{N/2}(z) (Z[N/2] = Z[0]),
// E[k] = (Z[k] + conj Z[N/2-k]) / 2 and O[k] = (Z[k] - conj Z[N/2-k]) / 2i are the
// spectra of the even and odd samples and X[k] = E[k] + exp(-2 pi i k / N) O[k].
// Only X[0..N/2] is formed: the Hilbert mask zeroes the rest.
//
// Output m is x[center + D m] (D the decimation).  H carries the shift by
// `center`, so with F = 2^fold_log2 dividing D the samples x[center + F j] are
// the (N/F)-point inverse DFT of the aliased spectrum V[j] = sum_q Y[j + q N/F]
// (Y = X H), and output m is its sample (D/F) m: the inverse transform is N/F
// points long instead of N.
//
// The bins are read by the first forward pass, Y (or V) is formed by the first
// inverse pass as it reads its inputs, and the last inverse pass writes the
// demodulated output: no pass of its own for any of the three.  The first
// inverse pass reads every input it needs from the forward result before the
// workgroup barrier and writes after it, so it runs in place.  The buffer is
// conv_buffer_len(p) points (N/2 for real bins folded by F >= 2), the twiddle
// table follows it.
template <int NT, int NMAX, bool COMPLEX_BINS, int RMAX, bool W14 = false>
__device__ __forceinline__ void conv_line(const ConvParams& p, int line, float2* sbuf) {
#pragma clang fp contract(fast)
    const int N = p.fft_len;
    const int half = N / 2;
    float2* buf = sbuf;
    float2* twc = sbuf + conv_buffer_len(p);
    const int n_coarse = N >> p.twiddle_bits;
    const int n_fine = 1 << p.twiddle_bits;
    float2* twf = twc + n_coarse;
    for (int i = threadIdx.x; i < n_coarse + n_fine; i += NT) twc[i] = p.twiddles[i];
    __syncthreads();
    const Twiddle twN{twc, twf, p.twiddle_bits, n_fine - 1, 0};
    const int M = N >> p.fold_log2;  // inverse transform length
    const int log2m = p.log2_len - p.fold_log2;
    const Twiddle twM{twc, twf, p.twiddle_bits, n_fine - 1, p.fold_log2};
    const float2* __restrict__ H = p.H;
    // demodulated output: sample out_stride * m of the inverse transform is output m
    float2* out = p.iq + (size_t)line * p.n_out;
    const int n_out = p.n_out, E = p.out_stride;
    const float2* __restrict__ dem = p.demod;
    const auto store = [=](int o, float2 v) {
        if (E == 1) {
            if (o < n_out) out[o] = cmul(v, dem[o]);
        } else if (o % E == 0 && o / E < n_out) {
            out[o / E] = cmul(v, dem[o / E]);
        }
    };
    const bool real_in = !COMPLEX_BINS && !(p.noise_amplitude > 0.0f);
    if (real_in) {
        const float* src = static_cast<const float*>(p.bins) + (size_t)line * p.num_rf;
        const int nrf = p.num_rf;
        const auto load = [=](int m) {
            const int n = 2 * m;
            return make_float2(n < nrf ? src[n] : 0.f, n + 1 < nrf ? src[n + 1] : 0.f);
        };
        fft_any<NT, NMAX, false, RMAX, false, true>(buf, half, p.log2_len - 1,
                                                     Twiddle{twc, twf, p.twiddle_bits, n_fine - 1, 1}, load,
                                                     [buf](int i, float2 v) { buf[conv_slot(i)] = v; });
        // X[k] H[k] from Z[k] and Z[N/2 - k]
        const auto spec = [=](int k) -> float2 {
            const float2 a = buf[conv_slot(k & (half - 1))];
            const float2 c = buf[conv_slot((half - k) & (half - 1))];
            const float2 e = make_float2(0.5f * (a.x + c.x), 0.5f * (a.y - c.y));  // (a + conj c) / 2
            const float2 o = make_float2(0.5f * (a.y + c.y), 0.5f * (c.x - a.x));  // (a - conj c) / 2i
            return cmul(cadd(e, cmul(twN.template get<false>(k), o)), H[k]);
        };
        const auto fold = [=](int j) -> float2 {
            float2 acc = spec(j);
            for (int k = j + M; k <= half; k += M) acc = cadd(acc, spec(k));
            return acc;
        };
        const auto one_sided = [=](int k) -> float2 { return k <= half ? spec(k) : make_float2(0.f, 0.f); };
        if (p.fold_log2 == 0) fft_any<NT, NMAX, true, RMAX, true, false>(buf, M, log2m, twM, one_sided, store);
        else fft_any<NT, NMAX, true, RMAX, true, false>(buf, M, log2m, twM, fold, store);
    } else {
        // complex bins (phase delay) or noise: the full N-point transform
        const int nrf = p.num_rf;
        const float amp = p.noise_amplitude;
        const uint32_t seed = p.noise_seed;
        const uint32_t counter = p.noise_counter_ptr ? *p.noise_counter_ptr : p.noise_counter;
        const float* bins_r = static_cast<const float*>(p.bins);
        const float2* bins_c = static_cast<const float2*>(p.bins);
        const auto load = [=](int n) -> float2 {
            if (n >= nrf) return make_float2(0.f, 0.f);
            float2 v;
            if (COMPLEX_BINS) v = bins_c[(size_t)line * nrf + n];
            else v = make_float2(bins_r[(size_t)line * nrf + n], 0.f);
            if (amp > 0.0f) {
                const float2 z = normal_pair(seed, counter, (uint32_t)line, (uint32_t)n);
                v = make_float2(v.x + amp * z.x, v.y + amp * z.y);
            }
            return v;
        };
        fft_any<NT, NMAX, false, RMAX, false, true, W14>(buf, N, p.log2_len, twN, load,
                                                     [buf](int i, float2 v) { buf[conv_slot(i)] = v; });
        const auto fold = [=](int j) -> float2 {
            float2 acc = make_float2(0.f, 0.f);
            for (int k = j; k <= half; k += M) acc = cadd(acc, cmul(buf[conv_slot(k)], H[k]));
            return acc;
        };
        fft_any<NT, NMAX, true, RMAX, true, false>(buf, M, log2m, twM, fold, store);
    }
}

template <int NT, int NMAX, bool COMPLEX_BINS, int RMAX, bool W14 = false>
__global__ __launch_bounds__(NT) void convolve_demod_kernel(ConvParams p) {
    extern __shared__ float2 sbuf[];
    conv_line<NT, NMAX, COMPLEX_BINS, RMAX, W14>(p, blockIdx.x, sbuf);
}

// Convolution of the lines of many jobs of one variant: workgroup b takes item
// b, a (job, line) pair; items are ordered longest transforms first.  (A
// persistent loop over the items was measured to keep the passes' per-thread
// index arithmetic live across items: 128-160 VGPRs against 58-61 here.)
template <int NT, int NMAX, bool COMPLEX_BINS, int RMAX>
__global__ __launch_bounds__(NT) void convolve_batch_kernel(const ConvParams* __restrict__ jobs,
                                                            const int2* __restrict__ items) {
    extern __shared__ float2 sbuf[];
    const int2 w = items[blockIdx.x];
    conv_line<NT, NMAX, COMPLEX_BINS, RMAX>(jobs[w.x], w.y, sbuf);
}

template <int NT, int NMAX, bool CB, int R>
static hipError_t launch_conv_batch_t(const ConvParams* jobs, const int2* items, int n_items, size_t lds,
                                      hipStream_t stream) {
    auto k = convolve_batch_kernel<NT, NMAX, CB, R>;
    hipError_t e = allow_lds((const void*)k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(n_items), dim3(NT), lds, stream, jobs, items);
    return hipGetLastError();
}

hipError_t launch_convolve_batch(int variant, const ConvParams* jobs, const int2* items, int n_items, size_t lds,
                                 hipStream_t stream) {
    if (n_items == 0) return hipSuccess;
    const bool cb = variant & kConvComplex, r8 = variant & kConvRadix8;
#define BCSIM_CONVB(NT, NMAX)                                                                           \
    if (r8) return cb ? launch_conv_batch_t<NT, NMAX, true, 8>(jobs, items, n_items, lds, stream)       \
                      : launch_conv_batch_t<NT, NMAX, false, 8>(jobs, items, n_items, lds, stream);     \
    return cb ? launch_conv_batch_t<NT, NMAX, true, 4>(jobs, items, n_items, lds, stream)               \
              : launch_conv_batch_t<NT, NMAX, false, 4>(jobs, items, n_items, lds, stream)
    if (variant & kConvWide) {
        BCSIM_CONVB(1024, kMaxFftLen);
    }
    BCSIM_CONVB(256, 4096);
#undef BCSIM_CONVB
}

// Zero many regions, reset work counters and advance the jobs' noise frame
// counters in one launch (graph-safe: no memset nodes, and a replayed graph
// advances the noise stream like a fresh launch).
__global__ void fill_segments_kernel(const FillSeg* segs, int n_segs, int* counters, int n_counters,
                                     uint32_t* noise_counters, int n_noise) {
    if (blockIdx.x == 0 && (int)threadIdx.x < n_counters) counters[threadIdx.x] = 0;
    if (blockIdx.x == 1)
        for (int i = threadIdx.x; i < n_noise; i += blockDim.x) noise_counters[i] += 1u;
    for (int s = 0; s < n_segs; s++) {
        float* p = segs[s].p;
        const size_t n = segs[s].n;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
            p[i] = 0.0f;
    }
}

hipError_t launch_fill_segments(const FillSeg* segs, int n_segs, int* counters, int n_counters,
                                uint32_t* noise_counters, int n_noise, hipStream_t stream) {
    hipLaunchKernelGGL(fill_segments_kernel, dim3(1024), dim3(256), 0, stream, segs, n_segs, counters, n_counters,
                       noise_counters, n_noise);
    return hipGetLastError();
}

int conv_block_threads(int fft_len) { return fft_len >= 4096 ? 1024 : 256; }

template <int NT, int NMAX, bool CB, int R, bool W14 = false>
static hipError_t launch_conv(const ConvParams& p, size_t lds, hipStream_t stream) {
    const void* k = (const void*)convolve_demod_kernel<NT, NMAX, CB, R, W14>;
    hipError_t e = allow_lds(k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((convolve_demod_kernel<NT, NMAX, CB, R, W14>), dim3(p.num_lines), dim3(NT), lds, stream, p);
    return hipGetLastError();
}

hipError_t launch_convolve_demod(const ConvParams& p, hipStream_t stream) {
    if (p.num_lines == 0) return hipSuccess;
    const int nt = conv_block_threads(p.fft_len);
    const size_t lds = conv_lds_bytes(p);
    const bool cb = p.complex_bins != 0, r8 = p.radix == 8;
    if (nt == 1024) {
        if (cb && p.fft_len == 16384)  // the complex 2^14-point transform's own kernel (fft_any)
            return r8 ? launch_conv<1024, kMaxFftLen, true, 8, true>(p, lds, stream)
                      : launch_conv<1024, kMaxFftLen, true, 4, true>(p, lds, stream);
        if (r8) return cb ? launch_conv<1024, kMaxFftLen, true, 8>(p, lds, stream)
                          : launch_conv<1024, kMaxFftLen, false, 8>(p, lds, stream);
        return cb ? launch_conv<1024, kMaxFftLen, true, 4>(p, lds, stream)
                  : launch_conv<1024, kMaxFftLen, false, 4>(p, lds, stream);
    }
    if (r8) return cb ? launch_conv<256, 4096, true, 8>(p, lds, stream) : launch_conv<256, 4096, false, 8>(p, lds, stream);
    return cb ? launch_conv<256, 4096, true, 4>(p, lds, stream) : launch_conv<256, 4096, false, 4>(p, lds, stream);
}

}  // namespace bcsim_hip
