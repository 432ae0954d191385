/*
 * FP64 COLE oracle: TEST INFRASTRUCTURE ONLY.
 *
 * This is a CPU restatement, in plain C, of the reference's CPU algorithm
 * (src/core/algorithm/CpuAlgorithm.cpp:50-183,278-368) with every quantity
 * evaluated in double precision, except the radial bin index and the radial
 * distance it is computed from, which follow the reference's own float/double
 * arithmetic bit for bit (the "binning contract", see cole_oracle.c).  It is
 * the checker for the HIP path: only tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py may load it.  The product library never links
 * or calls it.
 *
 * The handle API mirrors the C-ABI of the product (include/bcsim_c.h) so the
 * same test driver can configure the product, this oracle and the compiled
 * reference identically.  Output is complex128, interleaved (re, im).
 */
#ifndef BCSIM_ORACLE_H
#define BCSIM_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

oracle_sim* oracle_create(void);
void        oracle_destroy(oracle_sim* s);
const char* oracle_last_error(const oracle_sim* s);

int oracle_set_parameter(oracle_sim* s, const char* key, const char* value);
int oracle_set_excitation(oracle_sim* s, const float* samples, int n, int center_index,
                          float fs, float demod_freq);
int oracle_set_scan_sequence(oracle_sim* s, int num_lines, const float* origins,
                             const float* directions, const float* lateral_dirs,
                             const float* timestamps, float line_length);
int oracle_clear_fixed_scatterers(oracle_sim* s);
int oracle_add_fixed_scatterers(oracle_sim* s, const float* xyza, size_t n);
int oracle_clear_spline_scatterers(oracle_sim* s);
int oracle_add_spline_scatterers(oracle_sim* s, int degree, const float* knots, int n_knots,
                                 const float* control_points, const float* amplitudes,
                                 size_t n, int num_cp);
int oracle_set_analytical_profile(oracle_sim* s, float sigma_lateral, float sigma_elevational);
int oracle_set_lut_profile(oracle_sim* s, float r_min, float r_max, float l_min, float l_max,
                           float e_min, float e_max, const float* samples,
                           int n_rad, int n_lat, int n_ele);
int oracle_get_output_shape(oracle_sim* s, int* num_lines, int* num_samples, int* num_rf,
                            int* fft_len);
/* IQ frame, [line][sample] complex128 interleaved; capacity in complex elements. */
int oracle_simulate_lines(oracle_sim* s, double* out, size_t capacity, int num_threads);
/* Time-projection bins before convolution, [line][num_rf] complex128. */
int oracle_project(oracle_sim* s, double* out, size_t capacity, int num_threads);

/* Building blocks, exported so tests can pin them one by one. */
size_t oracle_next_pow2(size_t n);
void   oracle_hilbert_mask(int n, double* out);
void   oracle_fft(double* data, int n, int inverse);  /* in place, complex interleaved, unscaled */
void   oracle_fft_conv(const double* x, int nx, const double* y, int ny, double* out);
int    oracle_num_rf_samples(float sound_speed, float line_length, float fs);
double oracle_bspline_basis(int j, int p, double t, const double* knots, int n_knots);
float  oracle_bspline_basis_f32(int j, int p, float t, const float* knots, int n_knots);
int    oracle_knot_interval(const float* knots, int n_knots, float t);
/* Binning contract: radial distance (float) and bin index for one point and line. */
int    oracle_contract_bin(const float* p, const float* origin, const float* dir,
                           const float* lat, int arc, float fs, float c, int spline_rule,
                           float* r_out);
double oracle_lut_sample(oracle_sim* s, double r, double l, double e);

#ifdef __cplusplus
}
#endif
#endif
