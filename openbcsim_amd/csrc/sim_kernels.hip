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

#include "proj_math.h"
#include "sim_device.h"

namespace bcsim_hip {

// ------------------------------------------------------------------------------------
// utilities

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
    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
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
    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
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
    // exp(-2 pi i idx / N) for 0 <= idx < N, conjugated for the inverse transform
    template <bool INV>
    __device__ __forceinline__ float2 get(int idx) const {
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

// One Stockham autosort pass of radix R over N points held in LDS.  Ls is the
// length of the sub-transforms already formed.  Every thread first reads all
// its inputs, the workgroup synchronises, then every thread writes its outputs,
// so one buffer suffices.
template <int R, int NT, int NMAX, bool INV>
__device__ __forceinline__ void stockham_pass(float2* buf, int N, int Ls, const Twiddle& tw) {
#pragma clang fp contract(fast)
    constexpr int IT = (NMAX / R + NT - 1) / NT;
    const int m = N / R;
    float2 x[IT][R];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int j = threadIdx.x + it * NT;
        if (j < m) {
#pragma unroll
            for (int q = 0; q < R; q++) x[it][q] = buf[j + q * m];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int j = threadIdx.x + it * NT;
        if (j < m) {
            const int k = j & (Ls - 1);
            const int step = k * (N / (Ls * R));
            if (R == 8) {
                // w^1, w^2, w^4 from the table; w^3, w^5, w^6, w^7 as products of two of
                // them: 6 table reads per butterfly instead of 14
                if (Ls > 1) {
                    const float2 w1 = tw.template get<INV>(step);
                    const float2 w2 = tw.template get<INV>(2 * step);
                    const float2 w4 = tw.template get<INV>(4 * step);
                    const float2 w3 = cmul(w1, w2);
                    x[it][1] = cmul(x[it][1], w1);
                    x[it][2] = cmul(x[it][2], w2);
                    x[it][3] = cmul(x[it][3], w3);
                    x[it][4] = cmul(x[it][4], w4);
                    x[it][5] = cmul(x[it][5], cmul(w1, w4));
                    x[it][6] = cmul(x[it][6], cmul(w2, w4));
                    x[it][7] = cmul(x[it][7], cmul(w3, w4));
                }
            } else if (Ls > 1) {  // Ls == 1: every twiddle is 1
#pragma unroll
                for (int q = 1; q < R; q++) x[it][q] = cmul(x[it][q], tw.template get<INV>(q * step));
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
            for (int s2 = 0; s2 < R; s2++) buf[base + s2 * Ls] = y[s2];
        }
    }
    __syncthreads();
}

// Power-of-two FFT in LDS: radix-RMAX passes (4 or 8) after one smaller pass
// that takes up the remainder of log2 N.
template <int NT, int NMAX, bool INV, int RMAX>
__device__ __forceinline__ void fft_lds(float2* buf, int N, int log2n, const Twiddle& tw) {
    int Ls = 1;
    if (RMAX == 8) {
        const int rem = log2n % 3;
        if (rem == 1) {
            stockham_pass<2, NT, NMAX, INV>(buf, N, Ls, tw);
            Ls = 2;
        } else if (rem == 2) {
            stockham_pass<4, NT, NMAX, INV>(buf, N, Ls, tw);
            Ls = 4;
        }
        while (Ls < N) {
            stockham_pass<8, NT, NMAX, INV>(buf, N, Ls, tw);
            Ls *= 8;
        }
    } else {
        if (log2n & 1) {
            stockham_pass<2, NT, NMAX, INV>(buf, N, Ls, tw);
            Ls = 2;
        }
        while (Ls < N) {
            stockham_pass<4, NT, NMAX, INV>(buf, N, Ls, tw);
            Ls *= 4;
        }
    }
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

// One line: bins -> FFT -> x H -> IFFT -> extract, demodulate, decimate.
template <int NT, int NMAX, bool COMPLEX_BINS, int RMAX>
__device__ __forceinline__ void conv_line(const ConvParams& p, int line, float2* sbuf) {
    const int N = p.fft_len;
    float2* buf = sbuf;
    float2* twc = sbuf + N;
    const int n_coarse = N >> p.twiddle_bits;
    const int n_fine = 1 << p.twiddle_bits;
    float2* twf = twc + n_coarse;
    for (int i = threadIdx.x; i < n_coarse + n_fine; i += NT) twc[i] = p.twiddles[i];
    // bins of this line, zero padded to N
    if (COMPLEX_BINS) {
        const float2* src = static_cast<const float2*>(p.bins) + (size_t)line * p.num_rf;
        for (int n = threadIdx.x; n < N; n += NT) buf[n] = n < p.num_rf ? src[n] : make_float2(0.f, 0.f);
    } else {
        const float* src = static_cast<const float*>(p.bins) + (size_t)line * p.num_rf;
        for (int n = threadIdx.x; n < N; n += NT) buf[n] = make_float2(n < p.num_rf ? src[n] : 0.f, 0.f);
    }
    if (p.noise_amplitude > 0.0f) {
        for (int n = threadIdx.x; n < p.num_rf; n += NT) {
            const float2 z = normal_pair(p.noise_seed, p.noise_counter, (uint32_t)line, (uint32_t)n);
            buf[n] = make_float2(buf[n].x + p.noise_amplitude * z.x, buf[n].y + p.noise_amplitude * z.y);
        }
    }
    __syncthreads();
    Twiddle tw{twc, twf, p.twiddle_bits, n_fine - 1};
    fft_lds<NT, NMAX, false, RMAX>(buf, N, p.log2_len, tw);
    // excitation spectrum with Hilbert mask and 1/N; zero above N/2 (the mask)
    const int half = N / 2;
    for (int k = threadIdx.x; k < N; k += NT) buf[k] = k <= half ? cmul(buf[k], p.H[k]) : make_float2(0.f, 0.f);
    __syncthreads();
    fft_lds<NT, NMAX, true, RMAX>(buf, N, p.log2_len, tw);
    // delay compensation, demodulation, decimation
    float2* out = p.iq + (size_t)line * p.n_out;
    for (int j = threadIdx.x; j < p.n_out; j += NT) out[j] = cmul(buf[p.center + j * p.decimation], p.demod[j]);
}

template <int NT, int NMAX, bool COMPLEX_BINS, int RMAX>
__global__ __launch_bounds__(NT) void convolve_demod_kernel(ConvParams p) {
    extern __shared__ float2 sbuf[];
    conv_line<NT, NMAX, COMPLEX_BINS, RMAX>(p, blockIdx.x, sbuf);
}

// Persistent convolution over the lines of many jobs: items are (job, line)
// pairs taken from a global counter, longest transforms first.
__global__ __launch_bounds__(1024) void convolve_batch_kernel(const ConvParams* __restrict__ jobs,
                                                              const int2* __restrict__ items, int n_items,
                                                              int* __restrict__ counter) {
    extern __shared__ float2 sbuf[];
    __shared__ int item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(counter, 1);
        __syncthreads();
        const int it = item_s;
        __syncthreads();
        if (it >= n_items) break;
        const int2 w = items[it];
        const ConvParams p = jobs[w.x];
        if (p.radix == 8) {
            if (p.complex_bins) conv_line<1024, kMaxFftLen, true, 8>(p, w.y, sbuf);
            else conv_line<1024, kMaxFftLen, false, 8>(p, w.y, sbuf);
        } else {
            if (p.complex_bins) conv_line<1024, kMaxFftLen, true, 4>(p, w.y, sbuf);
            else conv_line<1024, kMaxFftLen, false, 4>(p, w.y, sbuf);
        }
        __syncthreads();
    }
}

size_t convolve_batch_lds_bytes() {
    // the largest transform plus its two-level twiddle table (coarse and fine halves of log2 N)
    return ((size_t)kMaxFftLen + 2 * 128) * sizeof(float2);
}

hipError_t launch_convolve_batch(const ConvParams* jobs, const int2* items, int n_items, int* counter,
                                 int workgroups, hipStream_t stream) {
    if (n_items == 0) return hipSuccess;
    const size_t lds = convolve_batch_lds_bytes();
    hipError_t e = hipFuncSetAttribute((const void*)convolve_batch_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(convolve_batch_kernel, dim3(workgroups), dim3(1024), lds, stream, jobs, items, n_items,
                       counter);
    return hipGetLastError();
}

// Zero many regions and reset work counters in one launch (graph-safe: no memset nodes).
__global__ void fill_segments_kernel(const FillSeg* segs, int n_segs, int* counters, int n_counters) {
    if (blockIdx.x == 0 && (int)threadIdx.x < n_counters) counters[threadIdx.x] = 0;
    for (int s = 0; s < n_segs; s++) {
        float* p = segs[s].p;
        const size_t n = segs[s].n;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
            p[i] = 0.0f;
    }
}

hipError_t launch_fill_segments(const FillSeg* segs, int n_segs, int* counters, int n_counters,
                                hipStream_t stream) {
    hipLaunchKernelGGL(fill_segments_kernel, dim3(1024), dim3(256), 0, stream, segs, n_segs, counters, n_counters);
    return hipGetLastError();
}

int conv_block_threads(int fft_len) { return fft_len >= 4096 ? 1024 : 256; }

template <int NT, int NMAX, bool CB, int R>
static hipError_t launch_conv(const ConvParams& p, size_t lds, hipStream_t stream) {
    const void* k = (const void*)convolve_demod_kernel<NT, NMAX, CB, R>;
    hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((convolve_demod_kernel<NT, NMAX, CB, R>), dim3(p.num_lines), dim3(NT), lds, stream, p);
    return hipGetLastError();
}

hipError_t launch_convolve_demod(const ConvParams& p, hipStream_t stream) {
    if (p.num_lines == 0) return hipSuccess;
    const int nt = conv_block_threads(p.fft_len);
    const size_t lds = ((size_t)p.fft_len + (p.fft_len >> p.twiddle_bits) + (1u << p.twiddle_bits)) * sizeof(float2);
    const bool cb = p.complex_bins != 0, r8 = p.radix == 8;
    if (nt == 1024) {
        if (r8) return cb ? launch_conv<1024, kMaxFftLen, true, 8>(p, lds, stream)
                          : launch_conv<1024, kMaxFftLen, false, 8>(p, lds, stream);
        return cb ? launch_conv<1024, kMaxFftLen, true, 4>(p, lds, stream)
                  : launch_conv<1024, kMaxFftLen, false, 4>(p, lds, stream);
    }
    if (r8) return cb ? launch_conv<256, 4096, true, 8>(p, lds, stream) : launch_conv<256, 4096, false, 8>(p, lds, stream);
    return cb ? launch_conv<256, 4096, true, 4>(p, lds, stream) : launch_conv<256, 4096, false, 4>(p, lds, stream);
}

}  // namespace bcsim_hip
