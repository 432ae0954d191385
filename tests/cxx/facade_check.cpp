// A program written the way the reference's C++ users write them
// (src/examples/gpu_example1.cpp:47-177): bcsim::Create("gpu") and the
// IAlgorithm setters, here through the header-only facade include/bcsim/LibBCSim.hpp.
//
// Usage: facade_check <out.bin>
// Simulates a small fixed + spline phantom on a sector scan and writes the
// inputs and the IQ frame to out.bin, so a test can re-evaluate them with the
// FP64 oracle:
//   int32 L, S, N_fixed, N_spline, M, degree
//   float origins[L][3], dirs[L][3], lats[L][3], times[L]
//   float fixed[N_fixed][4], knots[M+degree+1], cps[N_spline][M][3], amps[N_spline]
//   float excitation[55]
//   complex64 iq[L][S]
#include <bcsim/LibBCSim.hpp>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>

namespace {
uint32_t g_state = 12345u;
float uniform(float a, float b) {  // a small LCG: enough for a fixed test phantom
    g_state = g_state * 1664525u + 1013904223u;
    return a + (b - a) * ((g_state >> 8) * (1.0f / 16777216.0f));
}
template <typename T>
void put(FILE* f, const T* p, size_t n) {
    if (fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("write failed");
}
}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: facade_check <out.bin>\n";
        return 2;
    }
    try {
        const int L = 24, NF = 3000, NS = 2000, M = 8, P = 3;
        const float fs = 50e6f, fc = 2.5e6f, line_length = 0.05f;
        auto sim = bcsim::Create("gpu");
        sim->set_parameter("sound_speed", "1540");
        sim->set_parameter("radial_decimation", "2");
        sim->set_parameter("phase_delay", "on");

        bcsim::ExcitationSignal ex;
        const int ne = 55;
        for (int i = 0; i < ne; i++) {
            const float t = (i - ne / 2) / fs;
            ex.samples.push_back(std::exp(-(t * t) / (2.0f * 1.6e-7f * 1.6e-7f)) * std::cos(6.2831853f * fc * t));
        }
        ex.center_index = ne / 2;
        ex.sampling_frequency = fs;
        ex.demod_freq = fc;
        sim->set_excitation(ex);

        auto seq = std::make_shared<bcsim::ScanSequence>(line_length);
        std::vector<float> o, d, l, ts;
        for (int i = 0; i < L; i++) {
            const double a = -0.5 + 1.0 * i / (L - 1);
            const bcsim::vector3 dir((float)std::sin(a), 0.0f, (float)std::cos(a));
            const bcsim::vector3 lat((float)std::cos(a), 0.0f, (float)-std::sin(a));
            const float t = uniform(0.0f, 0.99f);
            seq->add_scanline(bcsim::Scanline(bcsim::vector3(0, 0, 0), dir, lat, t));
            o.insert(o.end(), {0.f, 0.f, 0.f});
            d.insert(d.end(), {dir.x, dir.y, dir.z});
            l.insert(l.end(), {lat.x, lat.y, lat.z});
            ts.push_back(t);
        }
        sim->set_scan_sequence(seq);

        auto fixed = std::make_shared<bcsim::FixedScatterers>();
        std::vector<float> fx;
        for (int i = 0; i < NF; i++) {
            bcsim::PointScatterer p;
            p.pos = bcsim::vector3(uniform(-0.02f, 0.02f), uniform(-0.004f, 0.004f), uniform(0.005f, 0.05f));
            p.amplitude = uniform(-1.0f, 1.0f);
            fixed->scatterers.push_back(p);
            fx.insert(fx.end(), {p.pos.x, p.pos.y, p.pos.z, p.amplitude});
        }
        sim->add_fixed_scatterers(fixed);

        auto spl = std::make_shared<bcsim::SplineScatterers>();
        spl->spline_degree = P;
        for (int i = 0; i < P; i++) spl->knot_vector.push_back(0.0f);
        for (int i = 0; i <= M - P; i++) spl->knot_vector.push_back((float)i / (M - P));
        for (int i = 0; i < P; i++) spl->knot_vector.push_back(1.0f);
        std::vector<float> cps, amps;
        for (int s = 0; s < NS; s++) {
            const bcsim::vector3 base(uniform(-0.02f, 0.02f), uniform(-0.004f, 0.004f), uniform(0.005f, 0.05f));
            std::vector<bcsim::vector3> row;
            for (int j = 0; j < M; j++) {
                const bcsim::vector3 c = base + bcsim::vector3(uniform(-1e-3f, 1e-3f), uniform(-1e-3f, 1e-3f),
                                                               uniform(-1e-3f, 1e-3f));
                row.push_back(c);
                cps.insert(cps.end(), {c.x, c.y, c.z});
            }
            spl->control_points.push_back(row);
            spl->amplitudes.push_back(uniform(-1.0f, 1.0f));
        }
        amps = spl->amplitudes;
        sim->add_spline_scatterers(spl);
        sim->set_analytical_profile(std::make_shared<bcsim::GaussianBeamProfile>(1e-3f, 1.5e-3f));

        std::vector<std::vector<std::complex<float>>> lines;
        sim->simulate_lines(lines);
        if ((int)lines.size() != L) throw std::runtime_error("wrong number of lines");
        const int S = (int)lines[0].size();
        if (sim->get_total_num_scatterers() != (size_t)(NF + NS)) throw std::runtime_error("wrong scatterer count");

        FILE* f = fopen(argv[1], "wb");
        if (!f) throw std::runtime_error("cannot open output");
        const int32_t hdr[6] = {L, S, NF, NS, M, P};
        put(f, hdr, 6);
        put(f, o.data(), o.size());
        put(f, d.data(), d.size());
        put(f, l.data(), l.size());
        put(f, ts.data(), ts.size());
        put(f, fx.data(), fx.size());
        put(f, spl->knot_vector.data(), spl->knot_vector.size());
        put(f, cps.data(), cps.size());
        put(f, amps.data(), amps.size());
        put(f, ex.samples.data(), ex.samples.size());
        for (const auto& line : lines) put(f, reinterpret_cast<const float*>(line.data()), 2 * line.size());
        fclose(f);
        std::cout << "facade_check: " << L << " lines x " << S << " samples on "
                  << sim->get_parameter("cur_device_name") << std::endl;
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "facade_check failed: " << e.what() << std::endl;
        return 1;
    }
}
