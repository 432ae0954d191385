// Ceilings of the units that bound the scatterer projection, measured on the
// device by small kernels of the same access shape (measurement infrastructure
// of the benchmark; not part of the simulation path).
//
//   lds_add:   ds_add_f32 throughput.  Each workgroup of 1024 threads owns 4 rows
//              of 9740 float bins in LDS (a 4-line tile of the volume workload);
//              every wave adds, per step, one value per lane into one row, at
//              bins spread over a window of 256 (about the radial spread of one
//              chunk of 64 scatterers) whose start moves from step to step -- the
//              pattern of the tile kernel's additions.  mode 1: 64 consecutive
//              bins (one per bank), for comparison.  mode 2: the same pattern with
//              ds_add_u64 into 2 rows of 64-bit bins (the fixed-point bins).
//   gather:    32-byte record gathers of a LUT of nr x 64 x 64 records (two
//              16-byte loads each, as lut_weight reads a record), every wave's
//              64 lanes in a span_r x span_l x span_e cell neighbourhood that
//              moves to a random place every step: with a table that fits the L2
//              and a 5 x 10 x 10 span (64 scatterers of a chunk, ~3 mm apart, on
//              the volume workload's table), the rate of the projection's LUT
//              gathers without anything else in the way.
//   spline:    the control-point loads of a spline hit: every step, each wave
//              reads the 4 active control points (cubic) of 64 consecutive
//              splines at a random chunk and knot interval from the [cp][n] float
//              arrays x, y, z of an n x m dataset (12 coalesced 4-byte loads per
//              lane) and evaluates the position: spline hits of one wave per
//              second, the per-line spline projection's loads with nothing else
//              in the way.
//   valu:      independent FMA chains: VALU wave instructions per second.
//   copy:      a plain float4 copy (the calibration kernel of the HBM-traffic
//              counters: its bytes are known).
//
// C interface (libbcsim_ceilings.so, bound by openbcsim_amd/ceilings.py); each
// returns the kernel's duration in ms from HIP events on a private stream, or a
// negative value on a HIP error:
//   float ceil_lds_add(int workgroups, int steps, int mode, float* d_sink)
//   float ceil_gather(const float4* d_table, int nr, int span_r, int span_l, int span_e,
//                     int workgroups, int steps, float* d_sink)
//   float ceil_spline(const float* d_cx, const float* d_cy, const float* d_cz, int n, int m,
//                     int workgroups, int steps, float* d_sink)
//   float ceil_valu(int workgroups, int iters, float* d_sink)
//   float ceil_copy(float4* d_dst, const float4* d_src, long long n4)
//   int   ceil_cu_count(void)   (compute units of the current device, or -1)
#include <hip/hip_runtime.h>

namespace {

constexpr int kThreads = 1024;
constexpr int kRows = 4;
constexpr int kRowBins = 9740;
constexpr int kSpread = 256;  // power of two; 8192 + kSpread <= kRowBins

__device__ __forceinline__ unsigned hash32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// T float (ds_add_f32) or unsigned long long (ds_add_u64); ROWS rows of kRowBins
template <class T, int ROWS>
__global__ __launch_bounds__(kThreads) void lds_add_kernel(int steps, int mode, float* sink) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    T* bins = reinterpret_cast<T*>(lds_raw);
    for (int i = threadIdx.x; i < ROWS * kRowBins; i += kThreads) bins[i] = T(0);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    unsigned h = hash32((blockIdx.x * 16u + wave) * 0x9e3779b9u);
    const unsigned lane_off = mode == 1 ? (unsigned)lane : hash32(lane * 0x85ebca6bu + blockIdx.x);
    for (int s = 0; s < steps; s++) {
        h = hash32(h + s);  // wave-uniform: row and window start of this step
        const int row = h & (ROWS - 1);
        const int start = (h >> 2) & 8191;
        const int bin = start + (int)((lane_off + (unsigned)s * 37u * (mode != 1)) & (kSpread - 1));
        atomicAdd(&bins[row * kRowBins + bin], T(1));
    }
    __syncthreads();
    float acc = 0.0f;
    for (int i = threadIdx.x; i < ROWS * kRowBins; i += kThreads) acc += (float)bins[i];
    sink[blockIdx.x * kThreads + threadIdx.x] = acc;
}

// The gathers of a chunk's LUT lookups: a table of nr x 64 x 64 records of 32
// bytes; lane k of a wave keeps a fixed offset (dr_k, dl_k, de_k) in a
// span_r x span_l x span_e cell neighbourhood (the footprint of 64 scatterers a
// few millimetres apart on the table) and every step the wave's neighbourhood
// moves to a random cell.  Lanes share a 128-byte line only when their records
// fall in it, as in the projection.
__global__ __launch_bounds__(kThreads) void gather_kernel(const float4* __restrict__ table, int nr, int span_r,
                                                          int span_l, int span_e, int steps, float* sink) {
    constexpr int nl = 64, ne = 64;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const unsigned lh = hash32(lane * 0x27d4eb2du + 1u);
    const int off = (int)(lh % (unsigned)span_r) * nl * ne + (int)((lh >> 10) % (unsigned)span_l) * ne +
                    (int)((lh >> 20) % (unsigned)span_e);
    const unsigned rr = (unsigned)(nr - span_r + 1), ll = (unsigned)(nl - span_l + 1), ee = (unsigned)(ne - span_e + 1);
    unsigned h = hash32((blockIdx.x * 16u + wave) * 0x9e3779b9u);
    float acc = 0.0f;
    for (int s = 0; s < steps; s++) {
        h = hash32(h + s);  // wave-uniform: the neighbourhood of this step
        const int base = (int)(h % rr) * nl * ne + (int)((h >> 8) % ll) * ne + (int)((h >> 16) % ee);
        const int rec = base + off;
        const float4 a = table[2 * rec];
        const float4 b = table[2 * rec + 1];
        acc += (a.x + a.y) + (a.z + a.w) + (b.x + b.y) + (b.z + b.w);
    }
    sink[blockIdx.x * kThreads + threadIdx.x] = acc;
}

// A spline hit's control-point loads: a random chunk of 64 splines (n is a
// multiple of 64) and a random interval of 4 active control points per wave and
// step, position summed with fixed basis values.
__global__ __launch_bounds__(kThreads) void spline_kernel(const float* __restrict__ cx, const float* __restrict__ cy,
                                                          const float* __restrict__ cz, int n, int m, int steps,
                                                          float* sink) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    unsigned h = hash32((blockIdx.x * 16u + wave) * 0x9e3779b9u);
    const unsigned nchunks = (unsigned)(n / 64), intervals = (unsigned)(m - 3);
    float acc = 0.0f;
    for (int s = 0; s < steps; s++) {
        h = hash32(h + s);  // wave-uniform: chunk and interval of this step
        const size_t base = (size_t)((h >> 20) % intervals) * n + (size_t)(h % nchunks) * 64 + lane;
        float x[4], y[4], z[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            x[i] = cx[base + (size_t)i * n];
            y[i] = cy[base + (size_t)i * n];
            z[i] = cz[base + (size_t)i * n];
        }
        const float px = x[0] * 0.1f + x[1] * 0.4f + x[2] * 0.4f + x[3] * 0.1f;
        const float py = y[0] * 0.1f + y[1] * 0.4f + y[2] * 0.4f + y[3] * 0.1f;
        const float pz = z[0] * 0.1f + z[1] * 0.4f + z[2] * 0.4f + z[3] * 0.1f;
        acc += px * py + pz;
    }
    sink[blockIdx.x * kThreads + threadIdx.x] = acc;
}

// Independent FMA chains: the VALU's wave-instruction issue rate.
__global__ __launch_bounds__(kThreads) void valu_kernel(int iters, float* sink) {
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = (float)(threadIdx.x + k);
    const float x = 0.999f + 1e-6f * (float)blockIdx.x, y = 1e-3f;
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = fmaf(a[k], x, y);
    }
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) acc += a[k];
    sink[blockIdx.x * kThreads + threadIdx.x] = acc;
}

__global__ void copy_kernel(float4* __restrict__ dst, const float4* __restrict__ src, long long n4) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

template <class Launch>
float timed(Launch launch) {
    hipStream_t s = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    float ms = -1.0f;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return -1.0f;
    if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
        launch(s);  // warm-up: code object load, first-touch of the caches
        if (hipGetLastError() == hipSuccess && hipEventRecord(a, s) == hipSuccess) {
            launch(s);
            if (hipGetLastError() == hipSuccess && hipEventRecord(b, s) == hipSuccess &&
                hipEventSynchronize(b) == hipSuccess)
                (void)hipEventElapsedTime(&ms, a, b);
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    return ms;
}

}  // namespace

extern "C" {

__attribute__((visibility("default"))) float ceil_lds_add(int workgroups, int steps, int mode, float* sink) {
    if (mode < 0 || mode > 2) return -1.0f;
    const bool wide = mode == 2;
    const size_t lds = wide ? (size_t)2 * kRowBins * 8 : (size_t)kRows * kRowBins * sizeof(float);
    const void* k = wide ? (const void*)lds_add_kernel<unsigned long long, 2> : (const void*)lds_add_kernel<float, kRows>;
    if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return -1.0f;
    return timed([&](hipStream_t s) {
        if (wide)
            hipLaunchKernelGGL((lds_add_kernel<unsigned long long, 2>), dim3(workgroups), dim3(kThreads), lds, s, steps,
                               0, sink);
        else
            hipLaunchKernelGGL((lds_add_kernel<float, kRows>), dim3(workgroups), dim3(kThreads), lds, s, steps, mode,
                               sink);
    });
}

__attribute__((visibility("default"))) float ceil_gather(const float4* table, int nr, int span_r, int span_l,
                                                         int span_e, int workgroups, int steps, float* sink) {
    if (nr < 1 || span_r < 1 || span_l < 1 || span_e < 1 || span_r > nr || span_l > 64 || span_e > 64)
        return -1.0f;
    return timed([&](hipStream_t s) {
        hipLaunchKernelGGL(gather_kernel, dim3(workgroups), dim3(kThreads), 0, s, table, nr, span_r, span_l, span_e,
                           steps, sink);
    });
}

__attribute__((visibility("default"))) float ceil_spline(const float* cx, const float* cy, const float* cz, int n,
                                                         int m, int workgroups, int steps, float* sink) {
    if (n < 64 || n % 64 != 0 || m < 4) return -1.0f;
    return timed([&](hipStream_t s) {
        hipLaunchKernelGGL(spline_kernel, dim3(workgroups), dim3(kThreads), 0, s, cx, cy, cz, n, m, steps, sink);
    });
}

__attribute__((visibility("default"))) float ceil_valu(int workgroups, int iters, float* sink) {
    return timed([&](hipStream_t s) {
        hipLaunchKernelGGL(valu_kernel, dim3(workgroups), dim3(kThreads), 0, s, iters, sink);
    });
}

__attribute__((visibility("default"))) float ceil_copy(float4* dst, const float4* src, long long n4) {
    return timed([&](hipStream_t s) { hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, dst, src, n4); });
}

__attribute__((visibility("default"))) int ceil_cu_count() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    return cus;
}

}  // extern "C"
