/*
 * FP64 COLE oracle: TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates the reference CPU algorithm:
 *   - projection loops       src/core/algorithm/CpuAlgorithm.cpp:50-97 (fixed), :99-183 (spline)
 *   - per-line pipeline      src/core/algorithm/CpuAlgorithm.cpp:302-368
 *   - FFT convolver          src/core/BeamConvolver.cpp:41-104
 *   - Hilbert mask           src/core/discrete_hilbert_mask.hpp:37-66
 *   - FFT length rule        src/core/BeamConvolver.cpp:48-49, src/core/fft.cpp:80-82
 *   - RF sample count        src/core/algorithm/common_utils.hpp:37-41
 *   - B-spline basis         src/core/bspline.hpp:45-91
 *   - beam profiles          src/core/BeamProfile.cpp:49-130
 *   - scanline validation    src/core/ScanSequence.cpp:37-70
 *   - parameter defaults     src/core/algorithm/BaseAlgorithm.cpp:36-87
 *
 * Arithmetic.  Two things are NOT evaluated in double, because the reference
 * defines them by its own float arithmetic and both feed a discontinuous
 * function (the radial bin index):
 *   (1) the scatterer position of a spline at a line's time (float sum of
 *       float control points times float basis values, in index order), and
 *   (2) the beam-local coordinates r, l, e (float subtract, float dot products
 *       evaluated left to right, float sqrt, copysign).
 * The bin index is then floor(r*2*fs/c + 0.5) in double exactly as the
 * reference writes it.  This file must be compiled with -ffp-contract=off so
 * that (1) and (2) round exactly like the reference's x86-64 build.
 * Everything after that -- beam-profile weights, sub-sample phase, the bin
 * sums, the excitation spectrum, both FFTs, demodulation -- is double.
 *
 * Parameter values that the reference derives in float and then uses as
 * definitions are taken as the reference rounds them: the LUT grid spacing
 * (BeamProfile.cpp:68-70, a float quotient), the normalised demodulation
 * frequency f_demod/fs (CpuAlgorithm.cpp:354, a float quotient) and the
 * spline sampling step 1/fs (CpuAlgorithm.cpp:161).  The demodulation phase
 * itself is evaluated exactly, -2*pi*nfd*n in double; the reference rounds it
 * to float (CpuAlgorithm.cpp:355-357), which this oracle deliberately does not.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAX_DEGREE 4
#define PI_D 3.14159265358979323846

enum { PROFILE_NONE = 0, PROFILE_GAUSSIAN = 1, PROFILE_LUT = 2 };

typedef struct {
    float* xyza;
    size_t n;
} fixed_set;

typedef struct {
    int degree;
    int n_knots;
    int num_cp;
    size_t n;
    float* knots;
    float* cps;  /* [n][num_cp][3] */
    float* amps; /* [n] */
} spline_set;

typedef struct {
    float o[3], dir[3], lat[3], ele[3];
    float t;
} line_t;

struct oracle_sim {
    /* parameters (BaseAlgorithm.cpp:36-46 defaults) */
    float sound_speed;
    int arc;
    int phase_delay;
    int decimation;
    int sum_all_cs;
    double weight_cutoff; /* product parameter; 0 (the default) is the reference's sum */
    /* excitation */
    int have_exc;
    float* exc;
    int n_exc;
    int center;
    float fs;
    float f_demod;
    /* scan */
    int have_scan;
    int n_lines;
    line_t* lines;
    float line_length;
    /* scatterers */
    int n_fixed;
    fixed_set* fixed;
    int n_spline;
    spline_set* splines;
    /* profile */
    int profile;
    float sigma_l, sigma_e;
    float lut_r0, lut_r1, lut_l0, lut_l1, lut_e0, lut_e1;
    int lut_nr, lut_nl, lut_ne;
    float* lut;
    double lut_peak; /* max |sample|, the reference magnitude of the weight cutoff */
    char err[512];
};

static int fail(oracle_sim* s, const char* msg) {
    snprintf(s->err, sizeof(s->err), "%s", msg);
    return 1;
}

oracle_sim* oracle_create(void) {
    oracle_sim* s = (oracle_sim*)calloc(1, sizeof(oracle_sim));
    s->sound_speed = 1540.0f;
    s->arc = 1;
    s->phase_delay = 0;
    s->decimation = 1;
    return s;
}

void oracle_destroy(oracle_sim* s) {
    if (!s) return;
    oracle_clear_fixed_scatterers(s);
    oracle_clear_spline_scatterers(s);
    free(s->exc);
    free(s->lines);
    free(s->lut);
    free(s);
}

const char* oracle_last_error(const oracle_sim* s) { return s->err; }

static int parse_bool(const char* v, int* out) {
    if (!strcmp(v, "on") || !strcmp(v, "true")) { *out = 1; return 0; }
    if (!strcmp(v, "off") || !strcmp(v, "false")) { *out = 0; return 0; }
    return 1;
}

/* BaseAlgorithm.cpp:48-87 and CpuAlgorithm.cpp:223-251 (the keys this oracle needs). */
int oracle_set_parameter(oracle_sim* s, const char* key, const char* value) {
    if (!strcmp(key, "sound_speed")) {
        float c = strtof(value, NULL);
        if (!(c > 0)) return fail(s, "illegal sound speed");
        s->sound_speed = c;
    } else if (!strcmp(key, "use_arc_projection")) {
        if (parse_bool(value, &s->arc)) return fail(s, "invalid boolean value");
    } else if (!strcmp(key, "phase_delay")) {
        if (parse_bool(value, &s->phase_delay)) return fail(s, "invalid boolean value");
    } else if (!strcmp(key, "sum_all_cs")) {
        if (parse_bool(value, &s->sum_all_cs)) return fail(s, "invalid boolean value");
    } else if (!strcmp(key, "radial_decimation")) {
        int d = atoi(value);
        if (d <= 0) return fail(s, "illegal radial decimation value");
        s->decimation = d;
    } else if (!strcmp(key, "weight_cutoff")) {
        /* The product's cull of weights below this fraction of the profile's peak
         * (1 for the Gaussian, max |sample| of a LUT).  Not a reference parameter:
         * it lets the tests pin the product's rule; the default 0 keeps every pair. */
        const double c = strtod(value, NULL);
        if (!(c >= 0.0 && c < 1.0)) return fail(s, "weight_cutoff must lie in [0, 1)");
        s->weight_cutoff = c;
    } else if (!strcmp(key, "noise_amplitude")) {
        if (strtof(value, NULL) != 0.0f) return fail(s, "oracle evaluates noise-free frames only");
    } else if (!strcmp(key, "verbose")) {
        /* accepted, no effect */
    } else {
        return fail(s, "illegal parameter name");
    }
    return 0;
}

int oracle_set_excitation(oracle_sim* s, const float* samples, int n, int center_index,
                          float fs, float demod_freq) {
    if (n <= 0) return fail(s, "excitation must have at least one sample");
    if (center_index < 0 || center_index >= n) return fail(s, "center_index out of range");
    free(s->exc);
    s->exc = (float*)malloc(sizeof(float) * (size_t)n);
    memcpy(s->exc, samples, sizeof(float) * (size_t)n);
    s->n_exc = n;
    s->center = center_index;
    s->fs = fs;
    s->f_demod = demod_freq;
    s->have_exc = 1;
    return 0;
}

/* Vector3D<float>::dot, left to right (vector3.hpp:86-88). */
static float dot3f(const float* a, const float* b) {
    float t = a[0] * b[0];
    t = t + a[1] * b[1];
    t = t + a[2] * b[2];
    return t;
}

/* Scanline constructor and validity test (ScanSequence.cpp:37-70). */
int oracle_set_scan_sequence(oracle_sim* s, int num_lines, const float* origins,
                             const float* directions, const float* lateral_dirs,
                             const float* timestamps, float line_length) {
    if (!s->have_exc) return fail(s, "Excitation must be configured before scan sequence");
    if (num_lines < 1) return fail(s, "No scanlines in scansequence");
    line_t* lines = (line_t*)calloc((size_t)num_lines, sizeof(line_t));
    for (int i = 0; i < num_lines; i++) {
        line_t* L = &lines[i];
        for (int k = 0; k < 3; k++) {
            L->o[k] = origins[3 * i + k];
            L->dir[k] = directions[3 * i + k];
            L->lat[k] = lateral_dirs[3 * i + k];
        }
        L->t = timestamps[i];
        /* elevational = lateral x direction (vector3.hpp:96-101) */
        const float* a = L->lat;
        const float* b = L->dir;
        L->ele[0] = a[1] * b[2] - b[1] * a[2];
        L->ele[1] = b[0] * a[2] - a[0] * b[2];
        L->ele[2] = a[0] * b[1] - b[0] * a[1];
        const float th = 1e-4f;
        int ok = 1;
        if (fabsf(dot3f(L->dir, L->lat)) >= th) ok = 0;
        if (fabsf(dot3f(L->lat, L->ele)) >= th) ok = 0;
        if (fabsf(dot3f(L->ele, L->dir)) >= th) ok = 0;
        if (fabs((double)sqrtf(dot3f(L->dir, L->dir)) - 1.0) >= th) ok = 0;
        if (fabs((double)sqrtf(dot3f(L->lat, L->lat)) - 1.0) >= th) ok = 0;
        if (fabs((double)sqrtf(dot3f(L->ele, L->ele)) - 1.0) >= th) ok = 0;
        if (!ok) {
            free(lines);
            return fail(s, "Invalid Scanline: not orthonormal unit vectors");
        }
    }
    free(s->lines);
    s->lines = lines;
    s->n_lines = num_lines;
    s->line_length = line_length;
    s->have_scan = 1;
    return 0;
}

int oracle_clear_fixed_scatterers(oracle_sim* s) {
    for (int i = 0; i < s->n_fixed; i++) free(s->fixed[i].xyza);
    free(s->fixed);
    s->fixed = NULL;
    s->n_fixed = 0;
    return 0;
}

int oracle_add_fixed_scatterers(oracle_sim* s, const float* xyza, size_t n) {
    s->fixed = (fixed_set*)realloc(s->fixed, sizeof(fixed_set) * (size_t)(s->n_fixed + 1));
    fixed_set* f = &s->fixed[s->n_fixed++];
    f->n = n;
    f->xyza = (float*)malloc(sizeof(float) * 4 * (n ? n : 1));
    if (n) memcpy(f->xyza, xyza, sizeof(float) * 4 * n);
    return 0;
}

int oracle_clear_spline_scatterers(oracle_sim* s) {
    for (int i = 0; i < s->n_spline; i++) {
        free(s->splines[i].knots);
        free(s->splines[i].cps);
        free(s->splines[i].amps);
    }
    free(s->splines);
    s->splines = NULL;
    s->n_spline = 0;
    return 0;
}

int oracle_add_spline_scatterers(oracle_sim* s, int degree, const float* knots, int n_knots,
                                 const float* control_points, const float* amplitudes,
                                 size_t n, int num_cp) {
    if (degree < 0 || degree > MAX_DEGREE) return fail(s, "spline degree out of range");
    if (num_cp <= degree) return fail(s, "too few spline control points for given degree");
    if (n_knots != num_cp + degree + 1)
        return fail(s, "mismatch in number of nodes, degree and control points");
    s->splines = (spline_set*)realloc(s->splines, sizeof(spline_set) * (size_t)(s->n_spline + 1));
    spline_set* sp = &s->splines[s->n_spline++];
    sp->degree = degree;
    sp->n_knots = n_knots;
    sp->num_cp = num_cp;
    sp->n = n;
    sp->knots = (float*)malloc(sizeof(float) * (size_t)n_knots);
    memcpy(sp->knots, knots, sizeof(float) * (size_t)n_knots);
    size_t ncp = n * (size_t)num_cp * 3;
    sp->cps = (float*)malloc(sizeof(float) * (ncp ? ncp : 1));
    if (ncp) memcpy(sp->cps, control_points, sizeof(float) * ncp);
    sp->amps = (float*)malloc(sizeof(float) * (n ? n : 1));
    if (n) memcpy(sp->amps, amplitudes, sizeof(float) * n);
    return 0;
}

int oracle_set_analytical_profile(oracle_sim* s, float sigma_lateral, float sigma_elevational) {
    s->profile = PROFILE_GAUSSIAN;
    s->sigma_l = sigma_lateral;
    s->sigma_e = sigma_elevational;
    return 0;
}

/* LUTBeamProfile constructor (BeamProfile.cpp:54-72). */
int oracle_set_lut_profile(oracle_sim* s, float r_min, float r_max, float l_min, float l_max,
                           float e_min, float e_max, const float* samples,
                           int n_rad, int n_lat, int n_ele) {
    if (n_rad <= 1) return fail(s, "Too few radial samples");
    if (n_lat <= 1) return fail(s, "Too few lateral samples");
    if (n_ele <= 1) return fail(s, "Too few elevational samples");
    size_t tot = (size_t)n_rad * (size_t)n_lat * (size_t)n_ele;
    free(s->lut);
    s->lut = (float*)malloc(sizeof(float) * tot);
    memcpy(s->lut, samples, sizeof(float) * tot);
    s->lut_peak = 0.0;
    for (size_t i = 0; i < tot; i++) s->lut_peak = fmax(s->lut_peak, fabs((double)samples[i]));
    s->lut_r0 = r_min; s->lut_r1 = r_max;
    s->lut_l0 = l_min; s->lut_l1 = l_max;
    s->lut_e0 = e_min; s->lut_e1 = e_max;
    s->lut_nr = n_rad; s->lut_nl = n_lat; s->lut_ne = n_ele;
    s->profile = PROFILE_LUT;
    return 0;
}

/* ---------------------------------------------------------------- building blocks */

/* fft.cpp:80-82: smallest power of two >= n (the reference's pow/log form gives the
 * same value for every n up to 2^21, checked exhaustively by the tests). */
size_t oracle_next_pow2(size_t n) {
    size_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

/* discrete_hilbert_mask.hpp:37-66 */
void oracle_hilbert_mask(int n, double* out) {
    for (int i = 0; i < n; i++) {
        double k = 0.0;
        if (n % 2 == 0) {
            if (i == 0) k = 1.0;
            else if (i >= 1 && i <= n / 2 - 1) k = 2.0;
            else if (i == n / 2) k = 1.0;
        } else {
            if (i == 0) k = 1.0;
            else if (i >= 1 && i < (n + 1) / 2) k = 2.0;
        }
        out[i] = k;
    }
}

/* Iterative radix-2 DIT FFT in double, twiddles from cos/sin of the exact angle
 * (no recurrence).  Forward uses exp(-2*pi*i*k*n/N); inverse uses the + sign and
 * is unscaled.  Power-of-two n only. */
void oracle_fft(double* a, int n, int inverse) {
    if (n <= 1) return;
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = a[2 * i], ti = a[2 * i + 1];
            a[2 * i] = a[2 * j]; a[2 * i + 1] = a[2 * j + 1];
            a[2 * j] = tr; a[2 * j + 1] = ti;
        }
    }
    const double sgn = inverse ? 1.0 : -1.0;
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int k = 0; k < half; k++) {
            const double ang = sgn * 2.0 * PI_D * (double)k / (double)len;
            const double wr = cos(ang), wi = sin(ang);
            for (int i = k; i < n; i += len) {
                const int j = i + half;
                const double xr = a[2 * j] * wr - a[2 * j + 1] * wi;
                const double xi = a[2 * j] * wi + a[2 * j + 1] * wr;
                a[2 * j] = a[2 * i] - xr;
                a[2 * j + 1] = a[2 * i + 1] - xi;
                a[2 * i] += xr;
                a[2 * i + 1] += xi;
            }
        }
    }
}

/* fft.cpp:93-121: linear convolution of two real signals through a zero-padded FFT. */
void oracle_fft_conv(const double* x, int nx, const double* y, int ny, double* out) {
    const int nout = nx + ny - 1;
    const int n = (int)oracle_next_pow2((size_t)nout);
    double* a = (double*)calloc((size_t)(2 * n), sizeof(double));
    double* b = (double*)calloc((size_t)(2 * n), sizeof(double));
    for (int i = 0; i < nx; i++) a[2 * i] = x[i];
    for (int i = 0; i < ny; i++) b[2 * i] = y[i];
    oracle_fft(a, n, 0);
    oracle_fft(b, n, 0);
    for (int k = 0; k < n; k++) {
        const double re = a[2 * k] * b[2 * k] - a[2 * k + 1] * b[2 * k + 1];
        const double im = a[2 * k] * b[2 * k + 1] + a[2 * k + 1] * b[2 * k];
        a[2 * k] = re;
        a[2 * k + 1] = im;
    }
    oracle_fft(a, n, 1);
    for (int i = 0; i < nout; i++) out[i] = a[2 * i] / n;
    free(a);
    free(b);
}

/* common_utils.hpp:37-41 with T = float, as CpuAlgorithm.cpp:266 instantiates it. */
int oracle_num_rf_samples(float sound_speed, float line_length, float fs) {
    const float max_time = (float)(2.0 * (double)line_length / (double)sound_speed);
    const float prod = fs * max_time;
    return (int)floor((double)prod + 0.5);
}

/* bspline.hpp:45-70 in double (the oracle's own basis). */
static double special_div_d(double num, double den) {
    if (fabs(num) < 1e-6 && fabs(den) < 1e-6) return 0.0;
    return num / den;
}

double oracle_bspline_basis(int j, int p, double x, const double* knots, int n_knots) {
    (void)n_knots;
    if (p == 0) return (x >= knots[j] && x < knots[j + 1]) ? 1.0 : 0.0;
    const double left = special_div_d((x - knots[j]) * oracle_bspline_basis(j, p - 1, x, knots, n_knots),
                                      knots[j + p] - knots[j]);
    const double right = special_div_d((knots[j + 1 + p] - x) * oracle_bspline_basis(j + 1, p - 1, x, knots, n_knots),
                                       knots[j + 1 + p] - knots[j + 1]);
    return left + right;
}

/* bspline.hpp:45-70 with T = float: the basis values the reference actually sums with. */
static float special_div_f(float num, float den) {
    if (fabs((double)num) < 1e-6 && fabs((double)den) < 1e-6) return 0.0f;
    return num / den;
}

float oracle_bspline_basis_f32(int j, int p, float x, const float* knots, int n_knots) {
    (void)n_knots;
    if (p == 0) return (x >= knots[j] && x < knots[j + 1]) ? 1.0f : 0.0f;
    const float left = special_div_f((x - knots[j]) * oracle_bspline_basis_f32(j, p - 1, x, knots, n_knots),
                                     knots[j + p] - knots[j]);
    const float right = special_div_f((knots[j + 1 + p] - x) * oracle_bspline_basis_f32(j + 1, p - 1, x, knots, n_knots),
                                      knots[j + 1 + p] - knots[j + 1]);
    return left + right;
}

/* bspline.hpp:75-82; returns -1 when no interval contains t. */
int oracle_knot_interval(const float* knots, int n_knots, float t) {
    for (int i = 0; i < n_knots - 1; i++)
        if (t >= knots[i] && t < knots[i + 1]) return i;
    return -1;
}

/* ---------------------------------------------------------------- binning contract */

/* Beam-local coordinates in float, CpuAlgorithm.cpp:58-72 / :144-158. */
static void local_coords_f(const float* p, const line_t* L, int arc, float* r, float* l, float* e) {
    float d[3];
    d[0] = p[0] - L->o[0];
    d[1] = p[1] - L->o[1];
    d[2] = p[2] - L->o[2];
    float rr = dot3f(d, L->dir);
    *l = dot3f(d, L->lat);
    *e = dot3f(d, L->ele);
    if (arc) rr = copysignf(sqrtf(dot3f(d, d)), rr);
    *r = rr;
}

/* Fixed rule CpuAlgorithm.cpp:75; spline rule :161-162. */
static double bin_arg_fixed(float r, float fs, float c) {
    return (double)r * 2.0 * (double)fs / (double)c + (double)0.5f;
}
static double bin_arg_spline(float r, float fs, float c) {
    const float dt = (float)(1.0 / (double)fs);
    return (double)r * 2.0 / (double)(c * dt) + (double)0.5f;
}

int oracle_contract_bin(const float* p, const float* origin, const float* dir,
                        const float* lat, int arc, float fs, float c, int spline_rule,
                        float* r_out) {
    line_t L;
    memcpy(L.o, origin, sizeof(L.o));
    memcpy(L.dir, dir, sizeof(L.dir));
    memcpy(L.lat, lat, sizeof(L.lat));
    L.ele[0] = L.lat[1] * L.dir[2] - L.dir[1] * L.lat[2];
    L.ele[1] = L.dir[0] * L.lat[2] - L.lat[0] * L.dir[2];
    L.ele[2] = L.lat[0] * L.dir[1] - L.dir[0] * L.lat[1];
    float r, l, e;
    local_coords_f(p, &L, arc, &r, &l, &e);
    if (r_out) *r_out = r;
    const double arg = spline_rule ? bin_arg_spline(r, fs, c) : bin_arg_fixed(r, fs, c);
    if (!(arg >= -2147483648.0 && arg < 2147483648.0)) return -2147483647 - 1;
    return (int)floor(arg);
}

/* ---------------------------------------------------------------- beam profiles (double) */

/* LUTBeamProfile::sampleProfile (BeamProfile.cpp:79-130) in double; grid spacing as
 * the reference stores it (a float quotient, BeamProfile.cpp:68-70). */
double oracle_lut_sample(oracle_sim* s, double r, double l, double e) {
    const double dr = (double)((s->lut_r1 - s->lut_r0) / (float)(s->lut_nr - 1));
    const double dl = (double)((s->lut_l1 - s->lut_l0) / (float)(s->lut_nl - 1));
    const double de = (double)((s->lut_e1 - s->lut_e0) / (float)(s->lut_ne - 1));
    const double tr = (r - (double)s->lut_r0) / dr;
    const double tl = (l - (double)s->lut_l0) / dl;
    const double te = (e - (double)s->lut_e0) / de;
    if (!(fabs(tr) < 1e9 && fabs(tl) < 1e9 && fabs(te) < 1e9)) return 0.0;
    const int r0 = (int)tr, r1 = (int)(tr + 1);
    const int l0 = (int)tl, l1 = (int)(tl + 1);
    const int e0 = (int)te, e1 = (int)(te + 1);
    const double fr = tr - r0, fl = tl - l0, fe = te - e0;
    if (r0 < 0 || r0 >= s->lut_nr || r1 < 0 || r1 >= s->lut_nr) return 0.0;
    if (l0 < 0 || l0 >= s->lut_nl || l1 < 0 || l1 >= s->lut_nl) return 0.0;
    if (e0 < 0 || e0 >= s->lut_ne || e1 < 0 || e1 >= s->lut_ne) return 0.0;
    const int ne = s->lut_ne, nl = s->lut_nl;
#define LUTV(ir, il, ie) ((double)s->lut[(size_t)(ie) + (size_t)ne * (size_t)(il) + (size_t)nl * (size_t)ne * (size_t)(ir)])
    const double c000 = LUTV(r0, l0, e0), c001 = LUTV(r0, l0, e1);
    const double c010 = LUTV(r0, l1, e0), c011 = LUTV(r0, l1, e1);
    const double c100 = LUTV(r1, l0, e0), c101 = LUTV(r1, l0, e1);
    const double c110 = LUTV(r1, l1, e0), c111 = LUTV(r1, l1, e1);
#undef LUTV
    const double c00 = (1.0 - fr) * c000 + fr * c100;
    const double c10 = (1.0 - fr) * c010 + fr * c110;
    const double c01 = (1.0 - fr) * c001 + fr * c101;
    const double c11 = (1.0 - fr) * c011 + fr * c111;
    const double c0 = (1.0 - fl) * c00 + fl * c10;
    const double c1 = (1.0 - fl) * c01 + fl * c11;
    return (1.0 - fe) * c0 + fe * c1;
}

/* Beam-profile weight; with a weight cutoff c, weights below c times the
 * profile's peak magnitude (1 for the Gaussian, the LUT's max |sample|) are zero. */
static double profile_weight(oracle_sim* s, float r, float l, float e) {
    if (s->profile == PROFILE_GAUSSIAN) {
        /* BeamProfile.cpp:49-56 */
        const double tsl = 2.0 * (double)s->sigma_l * (double)s->sigma_l;
        const double tse = 2.0 * (double)s->sigma_e * (double)s->sigma_e;
        const double ld = l, ed = e;
        const double x = ld * ld / tsl + ed * ed / tse;
        if (s->weight_cutoff > 0.0 && x > -log(s->weight_cutoff)) return 0.0;
        return exp(-x);
    }
    const double w = oracle_lut_sample(s, (double)r, (double)l, (double)e);
    if (s->weight_cutoff > 0.0 && fabs(w) < s->weight_cutoff * s->lut_peak) return 0.0;
    return w;
}

/* ---------------------------------------------------------------- projection */

static void add_contribution(oracle_sim* s, double* bins, int num_rf, int bin, double true_index,
                             double w) {
    if (bin < 0 || bin >= num_rf) return;
    if (s->phase_delay) {
        /* CpuAlgorithm.cpp:85-92: phase of the sub-sample displacement */
        const double ss_delay = ((double)bin - true_index) / (double)s->fs;
        const double ph = 2.0 * PI_D * (double)s->f_demod * ss_delay;
        bins[2 * bin] += w * cos(ph);
        bins[2 * bin + 1] += w * sin(ph);
    } else {
        bins[2 * bin] += w;
    }
}

static void project_fixed(oracle_sim* s, const fixed_set* f, const line_t* L, double* bins, int num_rf) {
    for (size_t n = 0; n < f->n; n++) {
        const float* p = &f->xyza[4 * n];
        float r, l, e;
        local_coords_f(p, L, s->arc, &r, &l, &e);
        const double arg = bin_arg_fixed(r, s->fs, s->sound_speed);
        if (!(arg >= 0.0 && arg < (double)num_rf)) continue;
        const int bin = (int)floor(arg);
        const double true_index = (double)r * 2.0 * (double)s->fs / (double)s->sound_speed;
        const double w = profile_weight(s, r, l, e) * (double)p[3];
        add_contribution(s, bins, num_rf, bin, true_index, w);
    }
}

static int project_spline(oracle_sim* s, const spline_set* sp, const line_t* L, double* bins, int num_rf) {
    const int M = sp->num_cp, p = sp->degree;
    const int mu = oracle_knot_interval(sp->knots, sp->n_knots, L->t);
    if (mu < 0) return 1;
    int lo = mu - p, hi = mu;
    if (s->sum_all_cs) { lo = 0; hi = M - 1; }
    if (lo < 0 || hi >= M) return 2;
    float* basis = (float*)malloc(sizeof(float) * (size_t)M);
    for (int i = 0; i < M; i++)
        basis[i] = oracle_bspline_basis_f32(i, p, L->t, sp->knots, sp->n_knots);
    const float dt = (float)(1.0 / (double)s->fs);
    const double denom = (double)(s->sound_speed * dt);
    for (size_t n = 0; n < sp->n; n++) {
        /* CpuAlgorithm.cpp:137-140: float sum in index order */
        float pos[3] = {0.0f, 0.0f, 0.0f};
        const float* cp = &sp->cps[n * (size_t)M * 3];
        for (int i = lo; i <= hi; i++) {
            pos[0] = pos[0] + cp[3 * i + 0] * basis[i];
            pos[1] = pos[1] + cp[3 * i + 1] * basis[i];
            pos[2] = pos[2] + cp[3 * i + 2] * basis[i];
        }
        float r, l, e;
        local_coords_f(pos, L, s->arc, &r, &l, &e);
        const double arg = bin_arg_spline(r, s->fs, s->sound_speed);
        if (!(arg >= 0.0 && arg < (double)num_rf)) continue;
        const int bin = (int)floor(arg);
        const double true_index = (double)r * 2.0 / denom;
        const double w = profile_weight(s, r, l, e) * (double)sp->amps[n];
        add_contribution(s, bins, num_rf, bin, true_index, w);
    }
    free(basis);
    return 0;
}

static int check_ready(oracle_sim* s) {
    if (!s->have_scan) return fail(s, "Scan sequence not configured.");
    if (!s->have_exc) return fail(s, "Excitation not configured.");
    if (s->profile == PROFILE_NONE) return fail(s, "Beam profile not configured.");
    size_t total = 0;
    for (int i = 0; i < s->n_fixed; i++) total += s->fixed[i].n;
    for (int i = 0; i < s->n_spline; i++) total += s->splines[i].n;
    if (total == 0) return fail(s, "no scatterers are configured");
    for (int i = 0; i < s->n_spline; i++) {
        const spline_set* sp = &s->splines[i];
        for (int k = 0; k < s->n_lines; k++) {
            const int mu = oracle_knot_interval(sp->knots, sp->n_knots, s->lines[k].t);
            if (mu < 0) return fail(s, "compute_knot_interval : could not determine knot interval");
            if (!s->sum_all_cs && (mu - sp->degree < 0 || mu >= sp->num_cp))
                return fail(s, "timestamp outside the span of the spline basis");
        }
    }
    return 0;
}

int oracle_get_output_shape(oracle_sim* s, int* num_lines, int* num_samples, int* num_rf,
                            int* fft_len) {
    if (!s->have_scan || !s->have_exc) return fail(s, "not configured");
    const int nrf = oracle_num_rf_samples(s->sound_speed, s->line_length, s->fs);
    if (num_lines) *num_lines = s->n_lines;
    if (num_samples) *num_samples = (nrf + s->decimation - 1) / s->decimation;
    if (num_rf) *num_rf = nrf;
    if (fft_len) *fft_len = (int)oracle_next_pow2((size_t)(nrf + s->n_exc - 1));
    return 0;
}

static void project_line(oracle_sim* s, int line_no, double* bins, int num_rf) {
    const line_t* L = &s->lines[line_no];
    for (int i = 0; i < s->n_fixed; i++) project_fixed(s, &s->fixed[i], L, bins, num_rf);
    for (int i = 0; i < s->n_spline; i++) project_spline(s, &s->splines[i], L, bins, num_rf);
}

int oracle_project(oracle_sim* s, double* out, size_t capacity, int num_threads) {
    if (check_ready(s)) return 1;
    int L, S, nrf, N;
    oracle_get_output_shape(s, &L, &S, &nrf, &N);
    if (capacity < (size_t)L * (size_t)nrf) return fail(s, "output capacity too small");
    memset(out, 0, sizeof(double) * 2 * (size_t)L * (size_t)nrf);
#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel for schedule(dynamic)
#endif
    for (int k = 0; k < L; k++) project_line(s, k, out + 2 * (size_t)k * (size_t)nrf, nrf);
    (void)num_threads;
    return 0;
}

int oracle_simulate_lines(oracle_sim* s, double* out, size_t capacity, int num_threads) {
    if (check_ready(s)) return 1;
    int L, S, nrf, N;
    oracle_get_output_shape(s, &L, &S, &nrf, &N);
    if (capacity < (size_t)L * (size_t)S) return fail(s, "output capacity too small");

    /* excitation spectrum with Hilbert mask (BeamConvolver.cpp:77-91), 1/N folded in */
    double* H = (double*)calloc((size_t)(2 * N), sizeof(double));
    double* mask = (double*)malloc(sizeof(double) * (size_t)N);
    for (int i = 0; i < s->n_exc; i++) H[2 * i] = (double)s->exc[i];
    oracle_fft(H, N, 0);
    oracle_hilbert_mask(N, mask);
    for (int k = 0; k < N; k++) {
        H[2 * k] *= mask[k] / N;
        H[2 * k + 1] *= mask[k] / N;
    }
    free(mask);
    const double nfd = (double)(s->f_demod / s->fs);
    const int D = s->decimation;

#ifdef _OPENMP
    if (num_threads > 0) omp_set_num_threads(num_threads);
#pragma omp parallel
#endif
    {
        double* buf = (double*)malloc(sizeof(double) * 2 * (size_t)N);
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
        for (int k = 0; k < L; k++) {
            memset(buf, 0, sizeof(double) * 2 * (size_t)N);
            project_line(s, k, buf, nrf);
            oracle_fft(buf, N, 0);
            for (int q = 0; q < N; q++) {
                const double re = buf[2 * q] * H[2 * q] - buf[2 * q + 1] * H[2 * q + 1];
                const double im = buf[2 * q] * H[2 * q + 1] + buf[2 * q + 1] * H[2 * q];
                buf[2 * q] = re;
                buf[2 * q + 1] = im;
            }
            oracle_fft(buf, N, 1);
            double* o = out + 2 * (size_t)k * (size_t)S;
            for (int j = 0; j < S; j++) {
                /* BeamConvolver.cpp:70-72 delay extraction; CpuAlgorithm.cpp:353-365 demod+decimate */
                const int n = j * D;
                const double yr = buf[2 * (s->center + n)], yi = buf[2 * (s->center + n) + 1];
                double cyc = nfd * (double)n; /* exact: 24-bit by 31-bit product */
                cyc -= floor(cyc);
                const double ph = -2.0 * PI_D * cyc;
                const double c = cos(ph), sn = sin(ph);
                o[2 * j] = yr * c - yi * sn;
                o[2 * j + 1] = yr * sn + yi * c;
            }
        }
        free(buf);
    }
    free(H);
    (void)num_threads;
    return 0;
}
