// Driver for the compiled reference CPU algorithm: TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile (target `ref`) against the reference's own sources
// where they lie (src/core/*.cpp, src/core/algorithm/{Base,Cpu}Algorithm.cpp),
// into oracle/_ref/libbcsim_ref.so.  Nothing of the reference is copied: this
// file only includes its headers and calls bcsim::Create("cpu")
// (src/core/LibBCSim.cpp:38-48).  A subclass exposes the protected
// CpuAlgorithm::projection_loop (src/core/algorithm/CpuAlgorithm.hpp:88-92) so
// tests can read the time-projection bins the reference produces before its
// FFT convolution.
//
// The exported functions mirror oracle.h (prefix ref_), with complex64 output.
#include <complex>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "LibBCSim.hpp"
#include "algorithm/CpuAlgorithm.hpp"
#include "algorithm/common_utils.hpp"

namespace {

class ProbeAlgorithm : public bcsim::CpuAlgorithm {
public:
    // Bins of one line exactly as CpuAlgorithm::simulate_line fills them
    // (CpuAlgorithm.cpp:313-328), before noise and convolution.
    void project_line(int line_no, std::complex<float>* bins, size_t num_rf) {
        const auto& line = m_scan_sequence->get_scanline(line_no);
        for (auto& f : m_scatterers_collection.fixed_collections) projection_loop(f, line, bins, num_rf);
        for (auto& s : m_scatterers_collection.spline_collections) projection_loop(s, line, bins, num_rf);
    }
    size_t num_rf() const { return m_rf_line_num_samples; }
    int num_lines() const { return m_scan_sequence->get_num_lines(); }
    int decimation() const { return m_radial_decimation; }
};

struct RefSim {
    std::shared_ptr<ProbeAlgorithm> alg = std::make_shared<ProbeAlgorithm>();
    std::string err;
    float sound_speed = 1540.0f;
    float fs = 0.0f;
    int decimation = 1;
    float line_length = 0.0f;
    int num_lines = 0;
    bool have_scan = false;
};

template <typename F>
int guarded(RefSim* s, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        s->err = e.what();
        return 1;
    }
}

}  // namespace

extern "C" {

void* ref_create(void) { return new RefSim(); }
void ref_destroy(void* h) { delete static_cast<RefSim*>(h); }
const char* ref_last_error(void* h) { return static_cast<RefSim*>(h)->err.c_str(); }

int ref_set_parameter(void* h, const char* key, const char* value) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] {
        s->alg->set_parameter(key, value);
        if (std::string(key) == "sound_speed") s->sound_speed = std::stof(value);
        if (std::string(key) == "radial_decimation") s->decimation = std::stoi(value);
    });
}

int ref_set_excitation(void* h, const float* samples, int n, int center_index, float fs,
                       float demod_freq) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] {
        bcsim::ExcitationSignal ex;
        ex.samples.assign(samples, samples + n);
        ex.center_index = center_index;
        ex.sampling_frequency = fs;
        ex.demod_freq = demod_freq;
        s->alg->set_excitation(ex);
        s->fs = fs;
    });
}

int ref_set_scan_sequence(void* h, int num_lines, const float* origins, const float* directions,
                          const float* lateral_dirs, const float* timestamps, float line_length) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] {
        auto seq = std::make_shared<bcsim::ScanSequence>(line_length);
        for (int i = 0; i < num_lines; i++) {
            const bcsim::vector3 o(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]);
            const bcsim::vector3 d(directions[3 * i], directions[3 * i + 1], directions[3 * i + 2]);
            const bcsim::vector3 l(lateral_dirs[3 * i], lateral_dirs[3 * i + 1], lateral_dirs[3 * i + 2]);
            seq->add_scanline(bcsim::Scanline(o, d, l, timestamps[i]));
        }
        s->alg->set_scan_sequence(seq);
        s->line_length = line_length;
        s->num_lines = num_lines;
        s->have_scan = true;
    });
}

int ref_clear_fixed_scatterers(void* h) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] { s->alg->clear_fixed_scatterers(); });
}

int ref_add_fixed_scatterers(void* h, const float* xyza, size_t n) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] {
        auto f = std::make_shared<bcsim::FixedScatterers>();
        f->scatterers.resize(n);
        for (size_t i = 0; i < n; i++) {
            f->scatterers[i].pos = bcsim::vector3(xyza[4 * i], xyza[4 * i + 1], xyza[4 * i + 2]);
            f->scatterers[i].amplitude = xyza[4 * i + 3];
        }
        s->alg->add_fixed_scatterers(f);
    });
}

int ref_clear_spline_scatterers(void* h) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] { s->alg->clear_spline_scatterers(); });
}

int ref_add_spline_scatterers(void* h, int degree, const float* knots, int n_knots,
                              const float* control_points, const float* amplitudes, size_t n,
                              int num_cp) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] {
        auto sp = std::make_shared<bcsim::SplineScatterers>();
        sp->spline_degree = degree;
        sp->knot_vector.assign(knots, knots + n_knots);
        sp->control_points.resize(n);
        sp->amplitudes.assign(amplitudes, amplitudes + n);
        for (size_t i = 0; i < n; i++) {
            sp->control_points[i].resize(num_cp);
            for (int j = 0; j < num_cp; j++) {
                const float* c = control_points + (i * num_cp + j) * 3;
                sp->control_points[i][j] = bcsim::vector3(c[0], c[1], c[2]);
            }
        }
        s->alg->add_spline_scatterers(sp);
    });
}

int ref_set_analytical_profile(void* h, float sigma_lateral, float sigma_elevational) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] {
        s->alg->set_analytical_profile(
            std::make_shared<bcsim::GaussianBeamProfile>(sigma_lateral, sigma_elevational));
    });
}

int ref_set_lut_profile(void* h, float r_min, float r_max, float l_min, float l_max, float e_min,
                        float e_max, const float* samples, int n_rad, int n_lat, int n_ele) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] {
        auto lut = std::make_shared<bcsim::LUTBeamProfile>(
            n_rad, n_lat, n_ele, bcsim::Interval(r_min, r_max), bcsim::Interval(l_min, l_max),
            bcsim::Interval(e_min, e_max));
        for (int ir = 0; ir < n_rad; ir++)
            for (int il = 0; il < n_lat; il++)
                for (int ie = 0; ie < n_ele; ie++)
                    lut->setDiscreteSample(ir, il, ie,
                                           samples[(static_cast<size_t>(ir) * n_lat + il) * n_ele + ie]);
        s->alg->set_lookup_profile(lut);
    });
}

int ref_get_output_shape(void* h, int* num_lines, int* num_samples, int* num_rf) {
    auto* s = static_cast<RefSim*>(h);
    if (!s->have_scan) {
        s->err = "not configured";
        return 1;
    }
    const size_t nrf = bcsim::compute_num_rf_samples(s->sound_speed, s->line_length, s->fs);
    if (num_lines) *num_lines = s->num_lines;
    if (num_samples) *num_samples = static_cast<int>((nrf + s->decimation - 1) / s->decimation);
    if (num_rf) *num_rf = static_cast<int>(nrf);
    return 0;
}

// IQ frame from CpuAlgorithm::simulate_lines, [line][sample] complex64 interleaved.
int ref_simulate_lines(void* h, float* out, size_t capacity, int* num_lines, int* num_samples) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] {
        std::vector<std::vector<std::complex<float>>> lines;
        s->alg->simulate_lines(lines);
        const size_t L = lines.size();
        const size_t S = L ? lines[0].size() : 0;
        if (L * S > capacity) throw std::runtime_error("output capacity too small");
        for (size_t k = 0; k < L; k++)
            std::memcpy(out + 2 * k * S, lines[k].data(), sizeof(std::complex<float>) * S);
        if (num_lines) *num_lines = static_cast<int>(L);
        if (num_samples) *num_samples = static_cast<int>(S);
    });
}

// Time-projection bins, [line][num_rf] complex64 interleaved.
int ref_project(void* h, float* out, size_t capacity) {
    auto* s = static_cast<RefSim*>(h);
    return guarded(s, [&] {
        const size_t nrf = s->alg->num_rf();
        const int L = s->alg->num_lines();
        if (static_cast<size_t>(L) * nrf > capacity) throw std::runtime_error("output capacity too small");
        std::memset(out, 0, sizeof(float) * 2 * L * nrf);
        for (int k = 0; k < L; k++)
            s->alg->project_line(k, reinterpret_cast<std::complex<float>*>(out) + k * nrf, nrf);
    });
}

int ref_set_num_threads(void* h, int n) { return ref_set_parameter(h, "num_cpu_cores", std::to_string(n).c_str()); }

}  // extern "C"
