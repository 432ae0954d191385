// LibBCSim.hpp -- header-only C++ facade of the reference's public API over the
// C-ABI in bcsim_c.h.
//
// A program written against the reference's headers (src/core/LibBCSim.hpp,
// BCSimConfig.hpp, ScanSequence.hpp, BeamProfile.hpp) compiles against this
// header and runs on the MI355X path by linking libbcsim_hip.so:
//
//     auto sim = bcsim::Create("gpu");
//     sim->set_excitation(ex); sim->set_scan_sequence(seq); ...
//     sim->simulate_lines(lines);
//
// Types keep the reference's names, members and error behaviour (constructors
// validate like the reference and throw std::runtime_error; every failed C-ABI
// call is rethrown as std::runtime_error with the library's message).
#pragma once

#include <cmath>
#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../bcsim_c.h"

namespace bcsim {

// ---------------------------------------------------------------- BCSimConfig.hpp:50-165

template <typename T>
class Vector3D {
public:
    Vector3D() : x(T(0)), y(T(0)), z(T(0)) {}
    Vector3D(T x_, T y_, T z_) : x(x_), y(y_), z(z_) {}
    T norm() const { return std::sqrt(x * x + y * y + z * z); }
    T norm_squared() const { return x * x + y * y + z * z; }
    Vector3D& operator+=(const Vector3D& r) { x += r.x; y += r.y; z += r.z; return *this; }
    Vector3D& operator-=(const Vector3D& r) { x -= r.x; y -= r.y; z -= r.z; return *this; }
    Vector3D& operator*=(const T& s) { x *= s; y *= s; z *= s; return *this; }
    Vector3D& operator/=(const T& s) { x /= s; y /= s; z /= s; return *this; }
    Vector3D operator+(const Vector3D& r) const { return Vector3D(*this) += r; }
    Vector3D operator-(const Vector3D& r) const { return Vector3D(*this) -= r; }
    Vector3D operator*(const T& s) const { return Vector3D(*this) *= s; }
    Vector3D operator/(const T& s) const { return Vector3D(*this) /= s; }
    T dot(const Vector3D& r) const { return x * r.x + y * r.y + z * r.z; }
    void normalize() { *this /= norm(); }
    Vector3D cross(const Vector3D& v) const {
        return Vector3D(y * v.z - v.y * z, v.x * z - x * v.z, x * v.y - v.x * y);
    }
    T x, y, z;
};
typedef Vector3D<float> vector3;

class ILog {
public:
    typedef std::shared_ptr<ILog> ptr;
    enum LogType { DEBUG, INFO, WARNING, FATAL };
    virtual ~ILog() {}
    virtual void write(LogType type, const std::string& msg) = 0;
};

class DummyLog : public ILog {
    void write(LogType, const std::string&) override {}
};

struct Interval {
    Interval(float t0, float t1) : first(t0), last(t1) {}
    float first;
    float last;
};

struct ExcitationSignal {
    std::vector<float> samples;
    int center_index;
    float sampling_frequency;
    float demod_freq;
};

struct PointScatterer {
    vector3 pos;
    float amplitude;
};

struct Scatterers {
    typedef std::shared_ptr<Scatterers> s_ptr;
    virtual ~Scatterers() {}
    virtual int num_scatterers() const = 0;
};

struct FixedScatterers : public Scatterers {
    typedef std::shared_ptr<FixedScatterers> s_ptr;
    int num_scatterers() const override { return static_cast<int>(scatterers.size()); }
    std::vector<PointScatterer> scatterers;
};

struct SplineScatterers : public Scatterers {
    typedef std::shared_ptr<SplineScatterers> s_ptr;
    int num_scatterers() const override { return static_cast<int>(control_points.size()); }
    size_t get_num_control_points() const {
        if (num_scatterers() == 0) throw std::runtime_error("No scatterers in dataset");
        return control_points[0].size();
    }
    void get_time_limits(float& start_time, float& end_time) const {
        const int num_knots = static_cast<int>(knot_vector.size());
        const int num_cs = static_cast<int>(get_num_control_points());
        if (spline_degree >= num_knots || num_cs >= num_knots) throw std::logic_error("invalid spline configuration");
        start_time = knot_vector[spline_degree];
        end_time = knot_vector[num_cs] - 1e-5f;
    }
    int spline_degree;
    std::vector<float> knot_vector;
    std::vector<std::vector<vector3>> control_points;
    std::vector<float> amplitudes;
};

// ---------------------------------------------------------------- ScanSequence.hpp:39-123

class Scanline {
public:
    Scanline() : timestamp(0.0f) {}
    Scanline(const vector3& origin_, const vector3& direction_, const vector3& lateral_dir_, float timestamp_)
        : origin(origin_), direction(direction_), lateral_dir(lateral_dir_), timestamp(timestamp_) {
        elevational_dir = lateral_dir.cross(direction);
        if (!is_orthogonal()) throw std::runtime_error("Invalid Scanline: Not orthogonal unit vectors");
        if (!is_normalized()) throw std::runtime_error("Invalid Scanline: Not unit length vectors");
    }
    vector3 get_elevational_dir() const { return elevational_dir; }
    vector3 get_direction() const { return direction; }
    vector3 get_origin() const { return origin; }
    vector3 get_lateral_dir() const { return lateral_dir; }
    float get_timestamp() const { return timestamp; }
    bool is_valid() const { return is_orthogonal() && is_normalized(); }

private:
    bool is_orthogonal() const {
        const float th = 1e-4f;
        return std::abs(direction.dot(lateral_dir)) < th && std::abs(lateral_dir.dot(elevational_dir)) < th &&
               std::abs(elevational_dir.dot(direction)) < th;
    }
    bool is_normalized() const {
        const float th = 1e-4f;
        return std::abs(direction.norm() - 1.0) < th && std::abs(lateral_dir.norm() - 1.0) < th &&
               std::abs(elevational_dir.norm() - 1.0) < th;
    }
    vector3 origin, direction, lateral_dir;
    float timestamp;
    vector3 elevational_dir;
};

class ScanSequence {
public:
    typedef std::shared_ptr<ScanSequence> s_ptr;
    explicit ScanSequence(float line_length_) : line_length(line_length_), all_timestamps_equal(false) {}
    int get_num_lines() const { return static_cast<int>(scanlines.size()); }
    void add_scanline(Scanline l) { scanlines.push_back(l); }
    const Scanline& get_scanline(int index) const {
        if (index < 0 || index >= get_num_lines()) throw std::runtime_error("ScanSequence::Invalid index.");
        return scanlines[index];
    }
    bool is_valid() const {
        for (const auto& l : scanlines)
            if (!l.is_valid()) return false;
        return true;
    }
    float line_length;
    bool all_timestamps_equal;

private:
    std::vector<Scanline> scanlines;
};

// ---------------------------------------------------------------- BeamProfile.hpp:38-142

class IBeamProfile {
public:
    typedef std::shared_ptr<IBeamProfile> s_ptr;
    virtual ~IBeamProfile() {}
    virtual float sampleProfile(float r, float l, float e) = 0;
};

class GaussianBeamProfile : public IBeamProfile {
public:
    GaussianBeamProfile(float sigma_lateral, float sigma_elevational)
        : m_sigma_lateral(sigma_lateral), m_sigma_elevational(sigma_elevational) {}
    void setSigmaLateral(float s) { m_sigma_lateral = s; }
    void setSigmaElevational(float s) { m_sigma_elevational = s; }
    float sampleProfile(float, float l, float e) override {
        const float tsl = 2 * m_sigma_lateral * m_sigma_lateral;
        const float tse = 2 * m_sigma_elevational * m_sigma_elevational;
        return std::exp(-(l * l / tsl + e * e / tse));
    }
    float getSigmaLateral() const { return m_sigma_lateral; }
    float getSigmaElevational() const { return m_sigma_elevational; }

protected:
    float m_sigma_lateral;
    float m_sigma_elevational;
};

class LUTBeamProfile : public IBeamProfile {
public:
    LUTBeamProfile(int num_samples_rad, int num_samples_lat, int num_samples_ele, Interval range_range,
                   Interval lateral_range, Interval elevational_range)
        : m_nr(num_samples_rad), m_nl(num_samples_lat), m_ne(num_samples_ele), m_rr(range_range),
          m_lr(lateral_range), m_er(elevational_range) {
        if (m_nr <= 1) throw std::runtime_error("Too few radial samples");
        if (m_nl <= 1) throw std::runtime_error("Too few lateral samples");
        if (m_ne <= 1) throw std::runtime_error("Too few elevational samples");
        m_samples.assign(static_cast<size_t>(m_nr) * m_nl * m_ne, 0.0f);
        m_dr = (m_rr.last - m_rr.first) / (m_nr - 1);
        m_dl = (m_lr.last - m_lr.first) / (m_nl - 1);
        m_de = (m_er.last - m_er.first) / (m_ne - 1);
    }
    // Host-side evaluation with the reference's rule (BeamProfile.cpp:79-130).
    float sampleProfile(float r, float l, float e) override {
        const double tr = (r - m_rr.first) / m_dr, tl = (l - m_lr.first) / m_dl, te = (e - m_er.first) / m_de;
        const int r0 = static_cast<int>(tr), r1 = static_cast<int>(tr + 1);
        const int l0 = static_cast<int>(tl), l1 = static_cast<int>(tl + 1);
        const int e0 = static_cast<int>(te), e1 = static_cast<int>(te + 1);
        const float fr = static_cast<float>(tr - r0), fl = static_cast<float>(tl - l0), fe = static_cast<float>(te - e0);
        if (r0 < 0 || r0 >= m_nr || r1 < 0 || r1 >= m_nr) return 0.0f;
        if (l0 < 0 || l0 >= m_nl || l1 < 0 || l1 >= m_nl) return 0.0f;
        if (e0 < 0 || e0 >= m_ne || e1 < 0 || e1 >= m_ne) return 0.0f;
        auto at = [&](int ir, int il, int ie) { return m_samples[getIndex(ir, il, ie)]; };
        const float c00 = (1.0f - fr) * at(r0, l0, e0) + fr * at(r1, l0, e0);
        const float c10 = (1.0f - fr) * at(r0, l1, e0) + fr * at(r1, l1, e0);
        const float c01 = (1.0f - fr) * at(r0, l0, e1) + fr * at(r1, l0, e1);
        const float c11 = (1.0f - fr) * at(r0, l1, e1) + fr * at(r1, l1, e1);
        const float c0 = (1.0f - fl) * c00 + fl * c10;
        const float c1 = (1.0f - fl) * c01 + fl * c11;
        return (1.0f - fe) * c0 + fe * c1;
    }
    void setDiscreteSample(int ir, int il, int ie, float v) {
        if (ir < 0 || ir >= m_nr || il < 0 || il >= m_nl || ie < 0 || ie >= m_ne) return;
        m_samples[getIndex(ir, il, ie)] = v;
    }
    Interval getRangeRange() const { return m_rr; }
    Interval getLateralRange() const { return m_lr; }
    Interval getElevationalRange() const { return m_er; }
    int getNumSamplesRadial() const { return m_nr; }
    int getNumSamplesLateral() const { return m_nl; }
    int getNumSamplesElevational() const { return m_ne; }
    const std::vector<float>& samples() const { return m_samples; }

protected:
    long getIndex(int r, int l, int e) const { return e + m_ne * l + static_cast<long>(m_nl) * m_ne * r; }
    int m_nr, m_nl, m_ne;
    Interval m_rr, m_lr, m_er;
    double m_dr, m_dl, m_de;
    std::vector<float> m_samples;
};

// ---------------------------------------------------------------- LibBCSim.hpp:42-97

class IAlgorithm {
public:
    typedef std::shared_ptr<IAlgorithm> s_ptr;
    typedef std::unique_ptr<IAlgorithm> u_ptr;
    virtual ~IAlgorithm() {}
    virtual void set_parameter(const std::string&, const std::string& value) = 0;
    virtual std::string get_parameter(const std::string& key) const = 0;
    virtual void clear_fixed_scatterers() = 0;
    virtual void add_fixed_scatterers(FixedScatterers::s_ptr) = 0;
    virtual void clear_spline_scatterers() = 0;
    virtual void add_spline_scatterers(SplineScatterers::s_ptr) = 0;
    virtual void set_scan_sequence(ScanSequence::s_ptr new_scan_sequence) = 0;
    virtual void set_excitation(const ExcitationSignal& new_excitation) = 0;
    virtual void set_analytical_profile(IBeamProfile::s_ptr beam_profile) = 0;
    virtual void set_lookup_profile(IBeamProfile::s_ptr beam_profile) = 0;
    virtual void simulate_lines(std::vector<std::vector<std::complex<float>>>& rf_lines) = 0;
    virtual std::vector<double> get_debug_data(const std::string& identifier) const = 0;
    virtual size_t get_total_num_scatterers() const = 0;
    virtual void set_logger(ILog::ptr log_object) = 0;
};

namespace detail {

inline void check(bcsim_status s) {
    if (s != BCSIM_OK) throw std::runtime_error(bcsim_last_error());
}

// The "gpu" algorithm: every call goes through the C-ABI.
class HipAlgorithm : public IAlgorithm {
public:
    HipAlgorithm() { check(bcsim_create("gpu", &m_sim)); }
    ~HipAlgorithm() override { bcsim_destroy(m_sim); }
    HipAlgorithm(const HipAlgorithm&) = delete;
    HipAlgorithm& operator=(const HipAlgorithm&) = delete;

    void set_parameter(const std::string& key, const std::string& value) override {
        check(bcsim_set_parameter(m_sim, key.c_str(), value.c_str()));
    }
    std::string get_parameter(const std::string& key) const override {
        char buf[512];
        check(bcsim_get_parameter(m_sim, key.c_str(), buf, sizeof(buf)));
        return buf;
    }
    void clear_fixed_scatterers() override { check(bcsim_clear_fixed_scatterers(m_sim)); }
    void add_fixed_scatterers(FixedScatterers::s_ptr f) override {
        std::vector<float> xyza;
        xyza.reserve(f->scatterers.size() * 4);
        for (const auto& p : f->scatterers) {
            xyza.push_back(p.pos.x);
            xyza.push_back(p.pos.y);
            xyza.push_back(p.pos.z);
            xyza.push_back(p.amplitude);
        }
        check(bcsim_add_fixed_scatterers(m_sim, xyza.data(), f->scatterers.size()));
    }
    void clear_spline_scatterers() override { check(bcsim_clear_spline_scatterers(m_sim)); }
    void add_spline_scatterers(SplineScatterers::s_ptr s) override {
        const size_t n = s->control_points.size();
        const int m = n ? static_cast<int>(s->control_points[0].size()) : 0;
        std::vector<float> cps;
        cps.reserve(n * m * 3);
        for (const auto& row : s->control_points) {
            if (static_cast<int>(row.size()) != m) throw std::runtime_error("ragged control point arrays");
            for (const auto& c : row) {
                cps.push_back(c.x);
                cps.push_back(c.y);
                cps.push_back(c.z);
            }
        }
        if (s->amplitudes.size() != n) throw std::runtime_error("Mismatch between control_points and amplitudes");
        check(bcsim_add_spline_scatterers(m_sim, s->spline_degree, s->knot_vector.data(),
                                          static_cast<int>(s->knot_vector.size()), cps.data(), s->amplitudes.data(),
                                          n, m));
    }
    void set_scan_sequence(ScanSequence::s_ptr seq) override {
        const int L = seq->get_num_lines();
        std::vector<float> o(3 * L), d(3 * L), l(3 * L), t(L);
        for (int i = 0; i < L; i++) {
            const Scanline& sl = seq->get_scanline(i);
            const vector3 a = sl.get_origin(), b = sl.get_direction(), c = sl.get_lateral_dir();
            o[3 * i] = a.x; o[3 * i + 1] = a.y; o[3 * i + 2] = a.z;
            d[3 * i] = b.x; d[3 * i + 1] = b.y; d[3 * i + 2] = b.z;
            l[3 * i] = c.x; l[3 * i + 1] = c.y; l[3 * i + 2] = c.z;
            t[i] = sl.get_timestamp();
        }
        check(bcsim_set_scan_sequence(m_sim, L, o.data(), d.data(), l.data(), t.data(), seq->line_length,
                                      seq->all_timestamps_equal ? 1 : 0));
    }
    void set_excitation(const ExcitationSignal& ex) override {
        check(bcsim_set_excitation(m_sim, ex.samples.data(), static_cast<int>(ex.samples.size()), ex.center_index,
                                   ex.sampling_frequency, ex.demod_freq));
    }
    void set_analytical_profile(IBeamProfile::s_ptr p) override {
        auto g = std::dynamic_pointer_cast<GaussianBeamProfile>(p);
        if (!g) throw std::runtime_error("GpuAlgorithm: failed to cast beam profile");
        check(bcsim_set_analytical_profile(m_sim, g->getSigmaLateral(), g->getSigmaElevational()));
    }
    void set_lookup_profile(IBeamProfile::s_ptr p) override {
        auto lut = std::dynamic_pointer_cast<LUTBeamProfile>(p);
        if (!lut) throw std::runtime_error("GpuAlgorithm: failed to cast beam profile");
        const Interval r = lut->getRangeRange(), l = lut->getLateralRange(), e = lut->getElevationalRange();
        check(bcsim_set_lut_profile(m_sim, r.first, r.last, l.first, l.last, e.first, e.last,
                                    lut->samples().data(), lut->getNumSamplesRadial(),
                                    lut->getNumSamplesLateral(), lut->getNumSamplesElevational()));
    }
    void simulate_lines(std::vector<std::vector<std::complex<float>>>& rf_lines) override {
        int L = 0, S = 0;
        check(bcsim_get_output_shape(m_sim, &L, &S));
        std::vector<std::complex<float>> flat(static_cast<size_t>(L) * S);
        check(bcsim_simulate_lines(m_sim, reinterpret_cast<float*>(flat.data()), flat.size(), &L, &S));
        rf_lines.assign(L, std::vector<std::complex<float>>());
        for (int k = 0; k < L; k++) rf_lines[k].assign(flat.begin() + static_cast<size_t>(k) * S,
                                                       flat.begin() + static_cast<size_t>(k + 1) * S);
    }
    std::vector<double> get_debug_data(const std::string& id) const override {
        size_t n = 0;
        check(bcsim_get_debug_data(m_sim, id.c_str(), nullptr, 0, &n));
        std::vector<double> v(n);
        check(bcsim_get_debug_data(m_sim, id.c_str(), v.data(), v.size(), &n));
        return v;
    }
    size_t get_total_num_scatterers() const override {
        size_t n = 0;
        check(bcsim_get_total_num_scatterers(m_sim, &n));
        return n;
    }
    void set_logger(ILog::ptr log) override {
        m_log = log;
        if (!log) {
            check(bcsim_set_logger(m_sim, nullptr, nullptr));
            return;
        }
        check(bcsim_set_logger(m_sim, &HipAlgorithm::forward_log, this));
    }

private:
    static void forward_log(int level, const char* msg, void* user) {
        auto* self = static_cast<HipAlgorithm*>(user);
        if (self->m_log) self->m_log->write(static_cast<ILog::LogType>(level), msg);
    }
    bcsim_sim* m_sim = nullptr;
    ILog::ptr m_log;
};

}  // namespace detail

// Factory (LibBCSim.cpp:38-48).  "gpu" is the MI355X implementation.
inline IAlgorithm::s_ptr Create(const std::string& sim_type) {
    if (sim_type == "gpu") return std::make_shared<detail::HipAlgorithm>();
    throw std::runtime_error("Illegal algorithm type: " + sim_type);
}

}  // namespace bcsim
