// Test infrastructure (CPU only): drives the library through its C-ABI with the
// call pattern of the benchmark -- device-resident frames of a 4096-line volume,
// per-call and batched frames of uneven jobs with projection modes switched in
// between, and the serving loop (warm-up over every phantom/geometry pair,
// Poisson arrivals answered from a second thread, single simulators, bursts,
// per-call loops that change each simulator's scan, and so its frame shape,
// between frames) -- against the mock runtime of mock_hip.cpp, which aborts on
// any lifetime or extent error.
//   usage: host_logic_driver [scale] [poisson_requests] [rate_per_s] [with_batch]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>
#include <chrono>
#include <atomic>
#include <mutex>
#include <condition_variable>
#include <deque>
#include <unistd.h>

#include <hip/hip_runtime.h>

#include "../../include/bcsim_c.h"

#define CHK(x)                                                                          \
    do {                                                                                \
        bcsim_status s_ = (x);                                                          \
        if (s_ != BCSIM_OK) {                                                           \
            fprintf(stderr, "FAIL %s:%d %s -> %d %s\n", __FILE__, __LINE__, #x, s_, bcsim_last_error()); \
            exit(1);                                                                    \
        }                                                                               \
    } while (0)

static std::mt19937_64 rng(7);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
static double N01() { return std::normal_distribution<double>(0, 1)(rng); }

struct Scan {
    std::vector<float> o, d, l, t;
    float len;
    int L() const { return (int)t.size(); }
};

static Scan sector(int n, double width, double depth) {
    Scan s;
    s.len = (float)depth;
    for (int k = 0; k < n; k++) {
        double a = -0.5 * width + (n > 1 ? k * width / (n - 1) : 0);
        s.o.insert(s.o.end(), {0, 0, 0});
        s.d.insert(s.d.end(), {(float)sin(a), 0, (float)cos(a)});
        s.l.insert(s.l.end(), {(float)cos(a), 0, (float)-sin(a)});
        s.t.push_back(0);
    }
    return s;
}

static Scan pyramid(int naz, int nel, double haz, double hel, double depth) {
    Scan s;
    s.len = (float)depth;
    for (int i = 0; i < naz; i++)
        for (int j = 0; j < nel; j++) {
            double az = -haz + (naz > 1 ? 2 * haz * i / (naz - 1) : 0);
            double el = -hel + (nel > 1 ? 2 * hel * j / (nel - 1) : 0);
            s.o.insert(s.o.end(), {0, 0, 0});
            s.d.insert(s.d.end(), {(float)(cos(el) * sin(az)), (float)sin(el), (float)(cos(el) * cos(az))});
            s.l.insert(s.l.end(), {(float)cos(az), 0, (float)-sin(az)});
            s.t.push_back(0);
        }
    return s;
}

struct Phantom {
    std::string kind;
    bcsim_sim* sim = nullptr;
    bool lut = false;
    float lut_r1 = 0;
    bool volume = false;
};

static std::vector<float> excitation(int& center) {
    std::vector<float> e(55);
    for (int i = 0; i < 55; i++) {
        double t = (i - 27) / 50e6;
        e[i] = (float)(cos(2 * M_PI * 2.5e6 * t) * exp(-t * t / (2 * 2.5e-7 * 2.5e-7)));
    }
    center = 27;
    return e;
}

static void base(bcsim_sim* s, const Scan& sc, int dec) {
    int c;
    auto e = excitation(c);
    CHK(bcsim_set_parameter(s, "sound_speed", "1540"));
    CHK(bcsim_set_parameter(s, "radial_decimation", std::to_string(dec).c_str()));
    CHK(bcsim_set_excitation(s, e.data(), (int)e.size(), c, 50e6f, 2.5e6f));
    CHK(bcsim_set_scan_sequence(s, sc.L(), sc.o.data(), sc.d.data(), sc.l.data(), sc.t.data(), sc.len, 0));
}

static std::vector<float> lut_table(int nr, int nl, int ne) {
    std::vector<float> t((size_t)nr * nl * ne);
    for (size_t i = 0; i < t.size(); i++) t[i] = (float)U(0, 1);
    return t;
}

static void add_fixed(bcsim_sim* s, size_t n, double y) {
    std::vector<float> p(n * 4);
    for (size_t i = 0; i < n; i++) {
        p[4 * i] = (float)U(-0.03, 0.03);
        p[4 * i + 1] = (float)U(-y, y);
        p[4 * i + 2] = (float)U(0.01, 0.12);
        p[4 * i + 3] = (float)N01();
    }
    CHK(bcsim_add_fixed_scatterers(s, p.data(), n));
}

static void add_spline(bcsim_sim* s, size_t n, int m) {
    const int deg = 3;
    std::vector<float> knots;
    for (int i = 0; i < deg; i++) knots.push_back(0);
    for (int i = 0; i <= m - deg; i++) knots.push_back((float)i / (m - deg));
    for (int i = 0; i < deg; i++) knots.push_back(1);
    std::vector<float> cp(n * m * 3), amp(n);
    for (size_t i = 0; i < n; i++) {
        float bx = (float)U(-0.03, 0.03), by = (float)U(-0.005, 0.005), bz = (float)U(0.01, 0.12);
        for (int k = 0; k < m; k++) {
            cp[(i * m + k) * 3] = bx + (float)(0.5e-3 * N01());
            cp[(i * m + k) * 3 + 1] = by + (float)(0.5e-3 * N01());
            cp[(i * m + k) * 3 + 2] = bz + (float)(0.5e-3 * N01());
        }
        amp[i] = (float)N01();
    }
    CHK(bcsim_add_spline_scatterers(s, deg, knots.data(), (int)knots.size(), cp.data(), amp.data(), n, m));
}

static Phantom make_phantom(int kind, double scale) {
    Phantom p;
    CHK(bcsim_create("gpu", &p.sim));
    if (kind == 0) {  // spline M-mode
        Scan sc = sector(64, 70 * M_PI / 180, 0.12);
        for (auto& t : sc.t) t = (float)U(0, 0.999);
        base(p.sim, sc, 4);
        add_spline(p.sim, (size_t)(200000 * scale), 10);
        CHK(bcsim_set_analytical_profile(p.sim, 1e-3f, 1e-3f));
        p.kind = "spline";
    } else if (kind == 1) {  // fixed 2D
        base(p.sim, sector(64, 70 * M_PI / 180, 0.12), 1);
        add_fixed(p.sim, (size_t)(100000 * scale), 0.005);
        CHK(bcsim_set_analytical_profile(p.sim, 1e-3f, 1e-3f));
        p.kind = "fixed";
    } else if (kind == 2) {  // volume LUT
        base(p.sim, pyramid(8, 8, M_PI / 6, M_PI / 6, 0.15), 2);
        add_fixed(p.sim, (size_t)(1000000 * scale), 0.03);
        auto t = lut_table(256, 64, 64);
        CHK(bcsim_set_lut_profile(p.sim, 0, 0.15f, -0.01f, 0.01f, -0.01f, 0.01f, t.data(), 256, 64, 64));
        p.kind = "volume";
        p.lut = true;
        p.lut_r1 = 0.15f;
        p.volume = true;
    } else {  // uneven job
        const double lens[4] = {0.028, 0.060, 0.12, 0.15};
        double len = lens[rng() % 4];
        size_t n = (size_t)(exp(U(log(10000.0), log(2000000.0))) * scale);
        if (n < 16) n = 16;
        int lines = 32 + (int)(rng() % 481);
        int dec = 1 << (rng() % 3);
        Scan sc = sector(lines, 70 * M_PI / 180, len);
        bool spl = rng() % 2;
        if (spl) for (auto& t : sc.t) t = (float)U(0, 0.999);
        base(p.sim, sc, dec);
        if (spl) add_spline(p.sim, n, 6 + (int)(rng() % 11));
        else add_fixed(p.sim, n, 0.005);
        if (rng() % 2) {
            auto t = lut_table(64, 32, 32);
            CHK(bcsim_set_lut_profile(p.sim, 0, (float)len, -0.01f, 0.01f, -0.01f, 0.01f, t.data(), 64, 32, 32));
            p.lut = true;
            p.lut_r1 = (float)len;
        } else {
            CHK(bcsim_set_analytical_profile(p.sim, 1e-3f, 1e-3f));
        }
        p.kind = "job";
    }
    return p;
}

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
    const double scale = argc > 1 ? atof(argv[1]) : 1.0;
    const int poisson_n = argc > 2 ? atoi(argv[2]) : 2000;
    const double rate = argc > 3 ? atof(argv[3]) : 200.0;
    const bool do_batch = argc > 4 ? atoi(argv[4]) != 0 : true;
    // light: fewer jobs, repetitions and requests (the sanitizer builds run 5-20x slower)
    const bool light = argc > 5 && atoi(argv[5]) != 0;
    const int n_jobs = light ? 16 : 64, n_sat = light ? 60 : 300, n_bursts = light ? 1 : 2;
    fprintf(stderr, "[driver] scale=%g poisson=%d rate=%g batch=%d light=%d\n", scale, poisson_n, rate, (int)do_batch,
            (int)light);

    // ---- one simulator whose scan length (and so its FFT length and output
    // size) changes between frames, with the output shape queried before each
    {
        Phantom f = make_phantom(1, scale);
        for (double len : {0.028, 0.12, 0.15, 0.06, 0.15, 0.028, 0.12}) {
            Scan sc = sector(48, 70 * M_PI / 180, len);
            CHK(bcsim_set_scan_sequence(f.sim, sc.L(), sc.o.data(), sc.d.data(), sc.l.data(), sc.t.data(), sc.len, 0));
            int L, S;
            CHK(bcsim_get_output_shape(f.sim, &L, &S));
            std::vector<float> host((size_t)L * S * 2);
            CHK(bcsim_simulate_lines(f.sim, host.data(), (size_t)L * S, &L, &S));
            // allocation-free frames after the tables exist (bins in the shared scratch)
            void* d_out = nullptr;
            hipMalloc(&d_out, (size_t)L * S * 8);
            for (int i = 0; i < 2; i++)
                CHK(bcsim_simulate_lines_device_noalloc(f.sim, d_out, (size_t)L * S, &L, &S));
            CHK(bcsim_synchronize(f.sim));
            hipFree(d_out);
        }
        CHK(bcsim_destroy(f.sim));
        fprintf(stderr, "[driver] scan-length changes done\n");
    }

    // ---- volume frames
    {
        Phantom v = make_phantom(2, scale);
        Scan sc = pyramid(64, 64, M_PI / 6, M_PI / 6, 0.15);
        CHK(bcsim_set_scan_sequence(v.sim, sc.L(), sc.o.data(), sc.d.data(), sc.l.data(), sc.t.data(), sc.len, 0));
        int L, S;
        CHK(bcsim_get_output_shape(v.sim, &L, &S));
        void* d_out = nullptr;
        hipMalloc(&d_out, (size_t)L * S * 8);
        for (int i = 0; i < 6; i++) CHK(bcsim_simulate_lines_device(v.sim, d_out, (size_t)L * S, &L, &S));
        CHK(bcsim_synchronize(v.sim));
        CHK(bcsim_set_parameter(v.sim, "store_kernel_details", "on"));
        for (int i = 0; i < 3; i++) CHK(bcsim_simulate_lines_device(v.sim, d_out, (size_t)L * S, &L, &S));
        CHK(bcsim_set_parameter(v.sim, "store_kernel_details", "off"));
        {
            // 4-line tiles of float bins leave the LDS room for 896 two-word list entries:
            // the plan holds the list as one-word entries instead
            char plan[512];
            CHK(bcsim_get_parameter(v.sim, "launch_plan", plan, sizeof plan));
            if (!std::strstr(plan, "lines=4") || !std::strstr(plan, "list=1792 packed")) {
                fprintf(stderr, "FAIL volume launch plan: %s\n", plan);
                exit(1);
            }
        }
        std::vector<float> host((size_t)L * S * 2);
        for (int i = 0; i < 3; i++) CHK(bcsim_simulate_lines(v.sim, host.data(), (size_t)L * S, &L, &S));
        CHK(bcsim_destroy(v.sim));
        hipFree(d_out);
        fprintf(stderr, "[driver] volume frames done\n");
    }

    // ---- 64 uneven jobs: per call in both projection modes, then batched
    if (do_batch) {
        std::vector<Phantom> jobs;
        for (int j = 0; j < n_jobs; j++) jobs.push_back(make_phantom(3, scale));
        bcsim_batch* b;
        CHK(bcsim_batch_create(&b));
        for (auto& p : jobs) CHK(bcsim_batch_add(b, p.sim, nullptr));
        size_t total;
        int nj;
        std::vector<size_t> offs(n_jobs);
        std::vector<int> ls(n_jobs), ss(n_jobs);
        CHK(bcsim_batch_get_layout(b, &total, &nj, offs.data(), ls.data(), ss.data(), n_jobs));
        void* d_out = nullptr;
        hipMalloc(&d_out, total * 8);
        for (const char* mode : {"direct", "tiled", "tiled"}) {
            for (auto& p : jobs) CHK(bcsim_set_parameter(p.sim, "projection_mode", mode));
            for (int rep = 0; rep < (light ? 1 : 3); rep++)
                for (int j = 0; j < n_jobs; j++) {
                    int L, S;
                    CHK(bcsim_simulate_lines_device(jobs[j].sim, (float2*)d_out + offs[j], (size_t)ls[j] * ss[j], &L, &S));
                    CHK(bcsim_synchronize(jobs[j].sim));
                }
        }
        for (int rep = 0; rep < 5; rep++) {
            CHK(bcsim_batch_simulate_device(b, d_out, total));
            CHK(bcsim_batch_synchronize(b));
        }
        CHK(bcsim_batch_destroy(b));
        for (auto& p : jobs) CHK(bcsim_destroy(p.sim));
        hipFree(d_out);
        fprintf(stderr, "[driver] batch done\n");
    }

    // ---- serving loop
    std::vector<Phantom> ph;
    for (int k = 0; k < 6; k++) ph.push_back(make_phantom(0, scale));
    for (int k = 0; k < 4; k++) ph.push_back(make_phantom(1, scale));
    for (int k = 0; k < 2; k++) ph.push_back(make_phantom(2, scale));
    for (int k = 0; k < 4; k++) ph.push_back(make_phantom(3, scale));
    std::vector<Scan> geo;
    for (int n : {32, 64, 128, 256, 512}) geo.push_back(sector(n, 70 * M_PI / 180, 0.12));
    geo.push_back(pyramid(64, 64, M_PI / 6, M_PI / 6, 0.15));
    std::vector<std::pair<int, int>> pairs;
    for (int p = 0; p < (int)ph.size(); p++)
        for (int g = 0; g < (int)geo.size(); g++) {
            if (ph[p].lut && geo[g].len > ph[p].lut_r1) continue;
            if (geo[g].L() == 4096 && !ph[p].volume) continue;
            pairs.push_back({p, g});
        }
    bcsim_server* srv;
    CHK(bcsim_server_create(&srv));
    for (auto& p : ph) CHK(bcsim_server_add_phantom(srv, p.sim, nullptr));
    for (auto& g : geo) CHK(bcsim_server_add_geometry(srv, g.L(), g.o.data(), g.d.data(), g.l.data(), g.len, nullptr));

    struct Req {
        uint64_t id;
        std::vector<float> out;
    };
    auto submit = [&](int p, int g, float t) {
        Req* r = new Req();
        int L, S;
        CHK(bcsim_server_get_output_shape(srv, p, g, &L, &S));
        r->out.resize((size_t)L * S * 2);
        CHK(bcsim_server_submit(srv, p, g, t, r->out.data(), (size_t)L * S, &r->id));
        return r;
    };
    auto wait = [&](Req* r) {
        CHK(bcsim_server_wait(srv, r->id, -1, nullptr, nullptr, nullptr));
        delete r;
    };
    {
        std::vector<Req*> w;
        for (auto& pg : pairs) w.push_back(submit(pg.first, pg.second, 0.5f));
        for (auto* r : w) wait(r);
        fprintf(stderr, "[driver] warmup done (%zu pairs)\n", pairs.size());
    }
    {
        std::deque<Req*> q;
        std::mutex mu;
        std::condition_variable cv;
        int done = 0;
        std::thread waiter([&] {
            for (int i = 0; i < poisson_n; i++) {
                Req* r;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return !q.empty(); });
                    r = q.front();
                    q.pop_front();
                }
                wait(r);
                done++;
            }
        });
        double t0 = now(), next = 0;
        for (int i = 0; i < poisson_n; i++) {
            next += std::exponential_distribution<double>(rate)(rng);
            while (now() - t0 < next) std::this_thread::sleep_for(std::chrono::microseconds(100));
            auto pg = pairs[rng() % pairs.size()];
            Req* r = submit(pg.first, pg.second, (float)U(0, 0.999));
            {
                std::lock_guard<std::mutex> g(mu);
                q.push_back(r);
            }
            cv.notify_one();
        }
        waiter.join();
        fprintf(stderr, "[driver] poisson done (%d)\n", done);
    }
    // sampled single simulators
    for (int k = 0; k < (light ? 3 : 10); k++) {
        auto pg = pairs[rng() % pairs.size()];
        bcsim_sim* s;
        CHK(bcsim_create("gpu", &s));
        Phantom& P = ph[pg.first];
        (void)P;
        // a fresh simulator of a new phantom of the same kind, with this geometry
        Phantom q = make_phantom(P.kind == "spline" ? 0 : P.kind == "fixed" ? 1 : P.kind == "volume" ? 2 : 3, scale);
        Scan sc = geo[pg.second];
        for (auto& t : sc.t) t = 0.3f;
        if (q.lut && sc.len > q.lut_r1) {
            CHK(bcsim_destroy(q.sim));
            CHK(bcsim_destroy(s));
            continue;
        }
        CHK(bcsim_set_scan_sequence(q.sim, sc.L(), sc.o.data(), sc.d.data(), sc.l.data(), sc.t.data(), sc.len, 1));
        int L, S;
        CHK(bcsim_get_output_shape(q.sim, &L, &S));
        std::vector<float> host((size_t)L * S * 2);
        CHK(bcsim_simulate_lines(q.sim, host.data(), (size_t)L * S, &L, &S));
        CHK(bcsim_destroy(q.sim));
        CHK(bcsim_destroy(s));
    }
    fprintf(stderr, "[driver] sampled singles done\n");
    // saturated bursts
    std::vector<std::tuple<int, int, float>> reqs;
    for (int i = 0; i < n_sat; i++) {
        auto pg = pairs[rng() % pairs.size()];
        reqs.push_back({pg.first, pg.second, (float)U(0, 0.999)});
    }
    for (int burst = 0; burst < n_bursts; burst++) {
        std::vector<Req*> w;
        for (auto& r : reqs) w.push_back(submit(std::get<0>(r), std::get<1>(r), std::get<2>(r)));
        for (auto* r : w) wait(r);
    }
    fprintf(stderr, "[driver] bursts done\n");
    // per-call loops on the phantom simulators
    for (const char* mode : {"tiled", "tiled", "direct"}) {
        for (auto& p : ph) CHK(bcsim_set_parameter(p.sim, "projection_mode", mode));
        for (auto& r : reqs) {
            Phantom& P = ph[std::get<0>(r)];
            Scan sc = geo[std::get<1>(r)];
            for (auto& t : sc.t) t = std::get<2>(r);
            CHK(bcsim_set_scan_sequence(P.sim, sc.L(), sc.o.data(), sc.d.data(), sc.l.data(), sc.t.data(), sc.len, 1));
            int L, S;
            CHK(bcsim_get_output_shape(P.sim, &L, &S));
            std::vector<float> host((size_t)L * S * 2);
            CHK(bcsim_simulate_lines(P.sim, host.data(), (size_t)L * S, &L, &S));
        }
    }
    fprintf(stderr, "[driver] per-call loops done\n");
    CHK(bcsim_server_destroy(srv));
    for (auto& p : ph) CHK(bcsim_destroy(p.sim));
    fprintf(stderr, "[driver] ALL DONE, no check failed\n");
    fflush(stderr);
    _exit(0);
}
