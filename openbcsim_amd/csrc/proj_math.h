// Device-side arithmetic of the scatterer projection, shared by the projection
// kernels.  Every function here reproduces the rounding of the reference's CPU
// code (CpuAlgorithm.cpp:50-183) where it selects a bin; translation units that
// include it are compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

#include "sim_device.h"

namespace bcsim_hip {

// Beam-local coordinates, float, left-to-right dot products (CpuAlgorithm.cpp:58-72).
struct Local {
    float r, l, e;
};

template <bool ARC>
__device__ __forceinline__ Local local_coords(const LineGeom& L, float px, float py, float pz) {
    const float dx = px - L.ox;
    const float dy = py - L.oy;
    const float dz = pz - L.oz;
    Local o;
    o.r = dx * L.rx + dy * L.ry + dz * L.rz;
    o.l = dx * L.lx + dy * L.ly + dz * L.lz;
    o.e = dx * L.ex + dy * L.ey + dz * L.ez;
    if (ARC) o.r = copysignf(__builtin_sqrtf(dx * dx + dy * dy + dz * dz), o.r);
    return o;
}

// Work done per thread, for the launch's statistics: profile-weight evaluations
// (LUT record gathers, Gaussian exponentials) and LDS bin additions.
struct Tally {
    int evals = 0;
    int adds = 0;
};

// GaussianBeamProfile::sampleProfile (BeamProfile.cpp:49-51); zero beyond the
// weight cutoff (exponent above g_cut).
__device__ __forceinline__ float gauss_weight(const ProfileParams& P, float l, float e, Tally& t) {
    const float x = l * l * P.inv_tsl + e * e * P.inv_tse;
    if (x > P.g_cut) return 0.0f;
    t.evals++;
    return expf(-x);
}

// LUTBeamProfile::sampleProfile (BeamProfile.cpp:79-130): truncating indices, zero
// whenever a neighbour falls outside the table, trilinear in r, then l, then e.
// The table is read as quad pairs: quad[2*((r*ne + e)*nl + l) + s] holds the
// samples at (l, e), (l, e+1), (l+1, e), (l+1, e+1) of radial slice r + s, so the
// eight corners are two 16-byte loads from one 32-byte record (one cache line).
// The lateral index runs fastest: lines next to each other in azimuth look up a
// point a few lateral cells apart, often in the same 128-byte cache line.
// Upper neighbour = lower + 1, except that a coordinate
// in (-1, 0) truncates to 0 for both (the reference's int(t) and int(t + 1)).
// The radial part depends on r only, so a caller projecting one point onto
// several lines from a common origin computes it once (lut_radial).  The weight
// feeds no discontinuity, so its interpolation uses fused lerps a + f*(b - a).
struct LutRadial {
    int off;   // cell offset of radial slice r0
    bool low;  // r in (-1, 0) steps: the upper radial neighbour is r0 itself
    float fr;
    bool ok;
    int4 box;  // with a weight cutoff: records of slice r0 that can reach it (ProfileParams::slice_box)
};

__device__ __forceinline__ LutRadial lut_radial(const ProfileParams& P, float r) {
    LutRadial R;
    const float tr = (r - P.r_min) * P.inv_dr;
    R.ok = tr > -1.0f && tr < (float)(P.nr - 1);
    const int r0 = R.ok ? (int)tr : 0;
    R.off = r0 * (P.nl * P.ne);
    R.low = tr < 0.0f;
    R.fr = tr - (float)r0;
    R.box = P.slice_box ? P.slice_box[r0] : make_int4(0, 0, 0, 0);
    return R;
}

__device__ __forceinline__ float lerp(float a, float b, float f) {
#pragma clang fp contract(fast)
    return a + f * (b - a);
}

__device__ __forceinline__ float lut_weight_le(const ProfileParams& P, const LutRadial& R, float l, float e,
                                               Tally& t) {
    const float tl = (l - P.l_min) * P.inv_dl;
    const float te = (e - P.e_min) * P.inv_de;
    if (!(R.ok && tl > -1.0f && tl < (float)(P.nl - 1) && te > -1.0f && te < (float)(P.ne - 1))) return 0.0f;
    const int l0 = (int)tl, e0 = (int)te;
    // weight cutoff: a record outside the slice's box holds only samples below it
    if (P.slice_box && (l0 < R.box.x || l0 > R.box.y || e0 < R.box.z || e0 > R.box.w)) return 0.0f;
    t.evals++;
    const float fl = tl - (float)l0, fe = te - (float)e0;
    const float4* cell = P.quad + 2 * (R.off + e0 * P.nl + l0);
    float4 a = cell[0];
    float4 b = cell[1];
    if (R.low) b = a;
    if (te < 0.0f || tl < 0.0f) {
        if (te < 0.0f) {
            a.y = a.x; a.w = a.z;
            b.y = b.x; b.w = b.z;
        }
        if (tl < 0.0f) {
            a.z = a.x; a.w = a.y;
            b.z = b.x; b.w = b.y;
        }
    }
    // radial, then lateral, then elevational (BeamProfile.cpp:117-129)
    const float c00 = lerp(a.x, b.x, R.fr);
    const float c10 = lerp(a.z, b.z, R.fr);
    const float c01 = lerp(a.y, b.y, R.fr);
    const float c11 = lerp(a.w, b.w, R.fr);
    const float c0 = lerp(c00, c10, fl);
    const float c1 = lerp(c01, c11, fl);
    const float w = lerp(c0, c1, fe);
    return fabsf(w) < P.w_cut ? 0.0f : w;  // w_cut is 0 without a weight cutoff
}

__device__ __forceinline__ float lut_weight(const ProfileParams& P, float r, float l, float e, Tally& t) {
    return lut_weight_le(P, lut_radial(P, r), l, e, t);
}

// Bin of radial distance r: floor(r*scale/denom + 0.5) evaluated in double as the
// reference writes it (CpuAlgorithm.cpp:75 fixed, :161-162 spline).  The quotient is
// first formed with the reciprocal; only when the result lies within 1e-9 of a
// bin boundary (where the two could round differently) is the exact division
// done.  Returns -1 for bins outside [0, num_rf) and for non-finite r.
__device__ __forceinline__ int radial_bin(const ProjCommon& c, float r, double& true_index) {
    const double q = (double)r * c.bin_scale;
    true_index = q * c.bin_inv_denom;
    double t = true_index + 0.5;
    double fl = floor(t);
    const double fr = t - fl;
    if (fr < 1e-9 || fr > 1.0 - 1e-9) {
        t = q / c.bin_denom + 0.5;
        fl = floor(t);
    }
    if (!(t >= 0.0 && t < (double)c.num_rf)) return -1;
    return (int)fl;
}

// Add one scatterer's contribution to its bin in LDS (CpuAlgorithm.cpp:77-95).
template <bool PHASE, bool LUT>
__device__ __forceinline__ void accumulate(const ProjCommon& c, float r, float l, float e, int bin,
                                           double true_index, float amp, float* lbins) {
    if (bin < 0) return;
    Tally t;
    const float w = (LUT ? lut_weight(c.prof, r, l, e, t) : gauss_weight(c.prof, l, e, t)) * amp;
    if (w == 0.0f) return;  // adding +-0 leaves the bin unchanged
    if (PHASE) {
        // CpuAlgorithm.cpp:85-92: w * exp(i * 2 pi f_demod * (bin - true_index) / fs)
        const float ph = (float)((double)bin - true_index) * c.phase_per_sample;
        float s, co;
        sincosf(ph, &s, &co);
        atomicAdd(&lbins[2 * bin], w * co);
        atomicAdd(&lbins[2 * bin + 1], w * s);
    } else {
        atomicAdd(&lbins[bin], w);
    }
}

// One (line, scatterer) pair from scratch: local coordinates, bin, weight.
template <bool ARC, bool PHASE, bool LUT>
__device__ __forceinline__ void project_point(const ProjCommon& c, const LineGeom& L, float px, float py,
                                              float pz, float amp, float* lbins) {
    const Local q = local_coords<ARC>(L, px, py, pz);
    double true_index;
    const int bin = radial_bin(c, q.r, true_index);
    accumulate<PHASE, LUT>(c, q.r, q.l, q.e, bin, true_index, amp, lbins);
}

__device__ __forceinline__ void zero_lds(float* lds, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = 0.0f;
}

// Write a tile of LDS bins to HBM: plain stores if this workgroup saw every
// scatterer of the frame, else an atomic add of the non-zero bins.
__device__ __forceinline__ void flush_bins(const float* lds, float* out, int n, bool atomic) {
    if (atomic) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const float v = lds[i];
            if (v != 0.0f) unsafeAtomicAdd(out + i, v);
        }
    } else {
        for (int i = threadIdx.x; i < n; i += blockDim.x) out[i] = lds[i];
    }
}

}  // namespace bcsim_hip
