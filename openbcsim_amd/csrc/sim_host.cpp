// Host side of the MI355X simulation path: the "gpu" algorithm behind the C-ABI
// of include/bcsim_c.h.
//
// Simulator plays the role of the reference's GpuAlgorithm
// (src/core/algorithm/GpuAlgorithm.cpp): it owns the device copies of the
// phantom, the scan and the beam profile, derives the per-frame tables, and
// enqueues a frame as
//     [zero bins] -> projection kernel per dataset -> fused convolve/demod kernel
// on one HIP stream, with no host round trip inside the frame (the reference
// issues about four launches and one constant upload per line, plus four
// device-wide synchronisations, GpuAlgorithm.cpp:189-360).
//
// Semantics follow the reference's CPU algorithm where the two reference
// implementations differ (FFT length next_pow2(num_rf + len - 1), bins limited
// to [0, num_rf), demodulation phase referenced to the extracted sample,
// trilinear LUT sampling on grid points); the error behaviour follows
// GpuAlgorithm (what it validates, and with which message).
#include "sim_state.h"

namespace bcsim_hip {
std::string& last_error() {
    thread_local std::string msg;
    return msg;
}
}  // namespace bcsim_hip


// -------------------------------------------------------------------------------------------
// C-ABI


extern "C" {

BCSIM_API const char* bcsim_last_error(void) { return bcsim_hip::last_error().c_str(); }

BCSIM_API bcsim_status bcsim_create(const char* sim_type, bcsim_sim** out) {
    return guarded([&] {
        if (!out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null output handle");
        *out = nullptr;
        const std::string t = sim_type ? sim_type : "";
        if (t != "gpu") fail(BCSIM_ERR_INVALID_ARGUMENT, "Illegal algorithm type: " + t);
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (count < 1) fail(BCSIM_ERR_DEVICE, "no HIP device available");
        *out = new bcsim_sim();
    });
}

BCSIM_API bcsim_status bcsim_create_simulator(const char* sim_type, bcsim_sim** out) {
    return bcsim_create(sim_type, out);
}

BCSIM_API bcsim_status bcsim_destroy(bcsim_sim* sim) {
    return guarded([&] { delete sim; });
}

BCSIM_API bcsim_status bcsim_set_parameter(bcsim_sim* sim, const char* key, const char* value) {
    return guarded([&] {
        need(sim);
        if (!key || !value) fail(BCSIM_ERR_INVALID_ARGUMENT, "null key or value");
        sim->set_parameter(key, value);
    });
}

BCSIM_API bcsim_status bcsim_get_parameter(bcsim_sim* sim, const char* key, char* buf, size_t buf_len) {
    return guarded([&] {
        need(sim);
        if (!key || !buf || buf_len == 0) fail(BCSIM_ERR_INVALID_ARGUMENT, "null key or buffer");
        const std::string v = sim->get_parameter(key);
        if (v.size() + 1 > buf_len) fail(BCSIM_ERR_CAPACITY, "buffer too small for parameter value");
        std::memcpy(buf, v.c_str(), v.size() + 1);
    });
}

BCSIM_API bcsim_status bcsim_set_excitation(bcsim_sim* sim, const float* samples, int n, int center_index,
                                            float fs, float demod_freq) {
    return guarded([&] {
        need(sim);
        sim->set_excitation(samples, n, center_index, fs, demod_freq);
    });
}

BCSIM_API bcsim_status bcsim_set_scan_sequence(bcsim_sim* sim, int num_lines, const float* origins,
                                               const float* directions, const float* lateral_dirs,
                                               const float* timestamps, float line_length,
                                               int all_timestamps_equal) {
    return guarded([&] {
        need(sim);
        sim->set_scan_sequence(num_lines, origins, directions, lateral_dirs, timestamps, line_length,
                               all_timestamps_equal != 0);
    });
}

BCSIM_API bcsim_status bcsim_clear_fixed_scatterers(bcsim_sim* sim) {
    return guarded([&] {
        need(sim);
        sim->clear_fixed();
    });
}

BCSIM_API bcsim_status bcsim_add_fixed_scatterers(bcsim_sim* sim, const float* xyza, size_t n) {
    return guarded([&] {
        need(sim);
        sim->add_fixed(xyza, n);
    });
}

BCSIM_API bcsim_status bcsim_set_fixed_scatterers(bcsim_sim* sim, const float* xyza, size_t n) {
    return guarded([&] {
        need(sim);
        sim->clear_fixed();
        sim->add_fixed(xyza, n);
    });
}

BCSIM_API bcsim_status bcsim_clear_spline_scatterers(bcsim_sim* sim) {
    return guarded([&] {
        need(sim);
        sim->clear_spline();
    });
}

BCSIM_API bcsim_status bcsim_add_spline_scatterers(bcsim_sim* sim, int degree, const float* knots, int n_knots,
                                                   const float* control_points, const float* amplitudes,
                                                   size_t n, int num_cp) {
    return guarded([&] {
        need(sim);
        sim->add_spline(degree, knots, n_knots, control_points, amplitudes, n, num_cp);
    });
}

BCSIM_API bcsim_status bcsim_set_spline_scatterers(bcsim_sim* sim, int degree, const float* knots, int n_knots,
                                                   const float* control_points, const float* amplitudes,
                                                   size_t n, int num_cp) {
    return guarded([&] {
        need(sim);
        sim->clear_spline();
        sim->add_spline(degree, knots, n_knots, control_points, amplitudes, n, num_cp);
    });
}

BCSIM_API bcsim_status bcsim_set_analytical_profile(bcsim_sim* sim, float sigma_lateral, float sigma_elevational) {
    return guarded([&] {
        need(sim);
        sim->set_gaussian(sigma_lateral, sigma_elevational);
    });
}

BCSIM_API bcsim_status bcsim_set_lut_profile(bcsim_sim* sim, float r_min, float r_max, float l_min, float l_max,
                                             float e_min, float e_max, const float* samples, int n_rad,
                                             int n_lat, int n_ele) {
    return guarded([&] {
        need(sim);
        sim->set_lut(r_min, r_max, l_min, l_max, e_min, e_max, samples, n_rad, n_lat, n_ele);
    });
}

BCSIM_API bcsim_status bcsim_get_output_shape(bcsim_sim* sim, int* num_lines, int* num_samples) {
    return guarded([&] {
        need(sim);
        sim->update_shape();
        if (num_lines) *num_lines = (int)sim->lines.size();
        if (num_samples) *num_samples = sim->n_out;
    });
}

BCSIM_API bcsim_status bcsim_simulate_lines(bcsim_sim* sim, float* out, size_t capacity, int* num_lines,
                                            int* num_samples) {
    return guarded([&] {
        need(sim);
        if (!out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null output buffer");
        sim->simulate_host(out, capacity, num_lines, num_samples);
    });
}

BCSIM_API bcsim_status bcsim_simulate_lines_device(bcsim_sim* sim, void* d_out, size_t capacity, int* num_lines,
                                                   int* num_samples) {
    return guarded([&] {
        need(sim);
        if (!d_out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null output buffer");
        sim->simulate_device(d_out, capacity, num_lines, num_samples);
    });
}

BCSIM_API bcsim_status bcsim_simulate_lines_device_noalloc(bcsim_sim* sim, void* d_out, size_t capacity,
                                                           int* num_lines, int* num_samples) {
    return guarded([&] {
        need(sim);
        if (!d_out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null output buffer");
        sim->simulate_device_noalloc(d_out, capacity, num_lines, num_samples);
    });
}

BCSIM_API bcsim_status bcsim_synchronize(bcsim_sim* sim) {
    return guarded([&] {
        need(sim);
        if (sim->stream) hip_check(hipStreamSynchronize(sim->stream), "hipStreamSynchronize");
    });
}

BCSIM_API bcsim_status bcsim_get_stream(bcsim_sim* sim, void** stream) {
    return guarded([&] {
        need(sim);
        if (!stream) fail(BCSIM_ERR_INVALID_ARGUMENT, "null output");
        sim->use_device();
        *stream = (void*)sim->stream;
    });
}

BCSIM_API bcsim_status bcsim_get_debug_data(bcsim_sim* sim, const char* key, double* out, size_t capacity,
                                            size_t* count) {
    return guarded([&] {
        need(sim);
        if (!key) fail(BCSIM_ERR_INVALID_ARGUMENT, "null key");
        auto it = sim->debug.find(key);
        if (it == sim->debug.end()) fail(BCSIM_ERR_INVALID_ARGUMENT, std::string("invalid debug data key: ") + key);
        if (count) *count = it->second.size();
        const size_t n = std::min(capacity, it->second.size());
        if (out && n) std::memcpy(out, it->second.data(), n * sizeof(double));
    });
}

BCSIM_API bcsim_status bcsim_get_total_num_scatterers(bcsim_sim* sim, size_t* out) {
    return guarded([&] {
        need(sim);
        if (!out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null output");
        *out = sim->total_fixed() + sim->total_spline();
    });
}

BCSIM_API bcsim_status bcsim_set_logger(bcsim_sim* sim, bcsim_log_fn fn, void* user) {
    return guarded([&] {
        need(sim);
        sim->log_fn = fn;
        sim->log_user = user;
    });
}

BCSIM_API bcsim_status bcsim_debug_get_projection(bcsim_sim* sim, float* out, size_t capacity, int* num_lines,
                                                  int* num_rf) {
    return guarded([&] {
        need(sim);
        sim->debug_projection(out, capacity, num_lines, num_rf);
    });
}

}  // extern "C"
