// Batched frames: many independent simulators (jobs) of uneven size simulated by
// one persistent projection launch and one persistent convolution launch.
//
// The reference has no batched path: every simulate_lines call is its own
// sequence of per-line launches (GpuAlgorithm.cpp:189-360).  Small jobs cannot
// fill 256 CUs on their own, so running many of them as separate calls leaves
// the GPU mostly idle; here each job keeps its own resident phantom, tables and
// bins, and the batch
//   1. zeroes the bins of jobs that accumulate and resets the work counters
//      (one fill kernel, no memset nodes, so the sequence can be captured),
//   2. renders all-timestamps-equal spline sets,
//   3. orders each job's blocks by the work they listed in the previous frame
//      (one order kernel, a workgroup per job) and runs one persistent
//      projection kernel over (job, tile-block) items, heaviest jobs first and
//      within a job its heaviest blocks first,
//   4. runs one persistent convolution kernel over (job, line) items, longest
//      transforms first, writing each job's IQ at its offset of the output.
// The sequence is captured once into a HIP graph and replayed while no job's
// configuration and the output pointer stay the same.
#pragma once
#include "sim_state.h"

#include <numeric>
#include <tuple>

namespace bcsim_hip {

struct JobPlan {
    bcsim_sim* sim;
    uint64_t version;
    size_t out_offset;  // complex elements
    int lines, samples;
};

}  // namespace bcsim_hip

struct bcsim_batch {
    int device = -1;
    hipStream_t stream = nullptr;
    bool use_graph = true;
    bool store_details = false;
    std::vector<bcsim_sim*> sims;

    // plan of the last build
    bool planned = false;
    std::vector<JobPlan> plan;
    size_t total_out = 0;
    DevBuf d_proj_jobs, d_proj_items, d_conv_jobs, d_conv_items, d_fill, d_counters, d_order_segs;
    // Per job, the noise frame counter of the frame in flight: set to the
    // simulator's counter minus one by build(), advanced by the fill kernel at the
    // start of every frame (graph replays included), so batched frames draw the
    // same noise as the simulator's own consecutive frames would.
    DevBuf d_noise;
    int n_proj_items = 0, n_conv_items = 0, n_fill = 0, n_order_segs = 0;
    // Projection jobs run as one persistent launch per kernel variant (JobFlags
    // that select code: arc, phase, LUT, spline, shared origin), each over its own
    // contiguous range of the item table with its own counter: a kernel holding
    // every variant needs more scalar registers than there are and spills them in
    // its inner loops.  Groups run heaviest first.
    struct ProjGroup {
        int variant, first, count;
        size_t lds;
    };
    std::vector<ProjGroup> groups;
    // the convolution items of one kernel variant (conv_variant): a launch each
    struct ConvGroup {
        int variant, first, count;
        size_t lds;  // the largest LDS any of its jobs needs
    };
    std::vector<ConvGroup> cgroups;
    struct Render {
        SplineSet set;
        SplineLineBasis basis;
        float4* out;
    };
    std::vector<Render> renders;
    int workgroups = 256;
    bool have_cu_count = false;
    bool autotune = true;          // tune each job's shape class before its first batched frame
    int tile_lines = 0;            // lines per tile for every job; 0 = each job's plan
    int items_per_workgroup = 4;   // projection items per persistent workgroup (balance vs. overhead)
    // Jobs of fewer blocks keep the index order (a segment of the order kernel each otherwise).
    static constexpr int kMinOrderedJobBlocks = 16;

    // graph of the last capture
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* graph_out = nullptr;

    std::map<std::string, std::vector<double>> debug;

    ~bcsim_batch() {
        drop_graph();
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (h_stage) (void)hipHostFree(h_stage);
    }

    // Device tables grow by at least half their size, so a batch whose mix of jobs
    // changes from frame to frame (the server's) soon stops reallocating: a
    // hipFree waits for every stream of the device.
    static void grow(DevBuf& d, size_t bytes) {
        if (bytes <= d.bytes && d.p) return;
        d.ensure(std::max<size_t>({bytes, d.bytes + d.bytes / 2, 256}));
    }

    // Pinned host staging of the tables, copied on `stream` without the host
    // waiting: a synchronous copy would wait for other streams' kernels that hold
    // the CUs.  Reused once the stream has drained (build() and set_output()
    // synchronise it first); the head [0, head) is kept for set_output's table.
    void* h_stage = nullptr;
    size_t h_stage_bytes = 0;
    bool just_built = false;
    void ensure_stage(size_t bytes) {
        if (bytes <= h_stage_bytes) return;
        if (h_stage) (void)hipHostFree(h_stage);
        h_stage = nullptr;
        h_stage_bytes = 0;
        const size_t n = std::max<size_t>(bytes, 64 << 10) * 2;
        hip_check(hipHostMalloc(&h_stage, n, 0), "hipHostMalloc");
        h_stage_bytes = n;
    }
    struct Stager {
        bcsim_batch& b;
        size_t head;
        std::vector<std::tuple<DevBuf*, const void*, size_t, size_t>> parts;  // device, source, bytes, offset
        size_t end;
        Stager(bcsim_batch& b_, size_t head_bytes) : b(b_), head((head_bytes + 255) & ~(size_t)255), end(head) {}
        void add(DevBuf& d, const void* src, size_t bytes) {
            parts.emplace_back(&d, src, bytes, end);
            end += (bytes + 255) & ~(size_t)255;
        }
        void upload() {
            b.ensure_stage(end);
            char* h = static_cast<char*>(b.h_stage);
            for (auto& p : parts) {
                const size_t bytes = std::get<2>(p);
                if (!bytes) continue;
                std::memcpy(h + std::get<3>(p), std::get<1>(p), bytes);
                hip_check(hipMemcpyAsync(std::get<0>(p)->p, h + std::get<3>(p), bytes, hipMemcpyHostToDevice, b.stream),
                          "hipMemcpyAsync");
            }
        }
    };

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_out = nullptr;
    }

    void invalidate() {
        planned = false;
        drop_graph();
    }

    bool stale() const {
        if (!planned || plan.size() != sims.size()) return true;
        for (size_t i = 0; i < sims.size(); i++)
            if (plan[i].sim != sims[i] || plan[i].version != sims[i]->version) return true;
        return false;
    }

    void add(bcsim_sim* s) {
        need(s);
        if (device < 0) device = s->device;
        if (s->device != device) fail(BCSIM_ERR_INVALID_ARGUMENT, "all simulators of a batch must use one device");
        sims.push_back(s);
        invalidate();
    }

    // Chunk slices per projection job for the persistent grid: items of about
    // equal estimated cost (chunks x lines, splines evaluated per line weighted
    // x3), `items_per_workgroup` per workgroup, so the heaviest-first order
    // balances without cutting light jobs into slices that only add atomics.
    void replan_slices(std::vector<ProjJob>& pjobs, const std::vector<double>& cost) const {
        double total = 0.0;
        for (double c : cost) total += c;
        const double per_item = total / ((double)workgroups * items_per_workgroup);
        for (size_t k = 0; k < pjobs.size(); k++) {
            ProjJob& J = pjobs[k];
            const double want_items = std::max(1.0, cost[k] / std::max(per_item, 1.0));
            int slices = (int)std::ceil(want_items / J.blocks_per_slice);
            slices = std::max(1, std::min(slices, std::max(1, J.nchunks / 16)));
            J.g.chunks_per_slice = std::max(1, (J.nchunks + slices - 1) / slices);
            J.g.slices = std::max(1, (J.nchunks + J.g.chunks_per_slice - 1) / J.g.chunks_per_slice);
        }
    }

    // Build the device-side job and item tables from the simulators' current state.
    void build() {
        if (sims.empty()) fail(BCSIM_ERR_INVALID_STATE, "empty batch");
        hip_check(hipSetDevice(device), "hipSetDevice");
        if (!stream) hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");  // frames in flight use the tables
        drop_graph();
        table_out = nullptr;
        if (!have_cu_count) {
            int cus = 0;
            hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
            workgroups = std::max(1, cus);
            have_cu_count = true;
        }
        plan.clear();
        renders.clear();
        std::vector<ProjJob> pjobs;
        std::vector<double> pcost;  // per-item cost estimate of each projection job
        std::vector<ConvParams> cjobs;
        std::vector<FillSeg> fills;
        std::vector<std::pair<int, int*>> ordered;  // (projection job, its work counts)
        std::vector<int> job_first_target, job_targets;
        std::vector<uint32_t> noise_init;
        size_t offset = 0;
        for (bcsim_sim* s : sims) {
            s->prepare();
            JobPlan jp{s, s->version, offset, (int)s->lines.size(), s->n_out};
            plan.push_back(jp);
            offset += (size_t)jp.lines * jp.samples;
            const bool cplx = s->phase_delay;
            const bool lut = s->profile == kProfileLut;
            const bool shared = s->shared_origin();
            const std::vector<bcsim_sim::ProjTarget> targets = s->proj_targets();
            const bool tune_now = autotune || s->autotune;
            for (size_t ti = 0; ti < targets.size(); ti++) {
                const auto& t = targets[ti];
                const bool is_spline = std::strcmp(t.kind, "spline") == 0;
                ProjJob J{};
                J.c = t.c;
                J.cp = s->cull_params(t.c);
                J.fs = t.fs;
                J.ss = t.ss;
                J.boxes = t.boxes;
                J.bins = s->d_bins.as<float>();
                J.g = s->plan_tile(t, tune_now);
                if (tile_lines > 0) {
                    TileTune tt = s->tune;
                    tt.lines = tile_lines;
                    J.g = s->tile_launch(t.nchunks, tt, is_spline, s->target_fixed(t));
                }
                J.g.block = 1024;
                J.g.chunks = t.chunk_order;
                J.g.chunk_dist = t.chunk_dist;
                J.nchunks = t.nchunks;
                J.flags = (s->arc ? kJobArc : 0) | (cplx ? kJobPhase : 0) | (lut ? kJobLut : 0) |
                          (is_spline ? kJobSpline : 0) | (shared && !is_spline ? kJobShared : 0) |
                          (J.g.fixed ? kJobFixed : 0);
                const int tiles = (t.c.num_lines + J.g.tile_lines - 1) / J.g.tile_lines;
                J.blocks_per_slice = tiles * J.g.windows;
                J.g.work = nullptr;
                J.g.order = nullptr;
                if (s->tile_order && J.blocks_per_slice >= kMinOrderedJobBlocks &&
                    J.blocks_per_slice <= kMaxOrderedTiles) {
                    // the job's blocks in the order of the work they listed in the previous
                    // frame, heaviest first (counts in the simulator's own buffer: zeroed on
                    // its stream, which quiesce() below waits for)
                    J.g.work = s->order_buffer(ti, J.blocks_per_slice);
                    ordered.push_back({(int)pjobs.size(), J.g.work});
                }
                pjobs.push_back(J);
                pcost.push_back((double)t.nchunks * t.c.num_lines * (is_spline ? 3.0 : 1.0));
                if (std::strcmp(t.kind, "rendered") == 0) {
                    Render r{t.ss, {}, t.scan->rendered.as<float4>()};
                    s->host_basis_line0(*t.sp, r.basis);
                    renders.push_back(r);
                }
            }
            job_first_target.push_back((int)pjobs.size() - (int)targets.size());
            job_targets.push_back((int)targets.size());
            const uint32_t ctr = s->noise_counter;
            cjobs.push_back(s->conv_params(nullptr));  // iq pointer set per call (offset into the output)
            s->noise_counter = ctr;  // a batched frame advances it when it is enqueued
            noise_init.push_back(ctr - 1u);
            cjobs.back().radix = s->plan_radix(tune_now);
            s->quiesce();  // renders and trials on the job's own stream are done before the batch runs
        }
        replan_slices(pjobs, pcost);
        // bins that are accumulated (or not written at all) start from zero
        for (size_t k = 0; k < sims.size(); k++) {
            bcsim_sim* s = sims[k];
            const int first = job_first_target[k], n = job_targets[k];
            bool single_store = n == 1 && pjobs[first].g.slices == 1;
            for (int q = first; q < first + n; q++) pjobs[q].g.atomic = single_store ? 0 : 1;
            if (!single_store)
                fills.push_back(FillSeg{s->d_bins.as<float>(),
                                        s->lines.size() * (size_t)s->num_rf * (s->phase_delay ? 2 : 1)});
        }
        total_out = offset;

        // items: grouped by kernel variant (groups heaviest first), jobs within a
        // group heaviest first
        std::vector<int> vkey(pjobs.size());
        std::map<int, double> group_cost;
        for (size_t j = 0; j < pjobs.size(); j++) {
            vkey[j] = batch_variant(pjobs[j].flags);
            group_cost[vkey[j]] += pcost[j];
        }
        std::vector<int> porder(pjobs.size());
        std::iota(porder.begin(), porder.end(), 0);
        std::stable_sort(porder.begin(), porder.end(), [&](int a, int b) {
            if (vkey[a] != vkey[b]) {
                const double ca = group_cost[vkey[a]], cb = group_cost[vkey[b]];
                return ca != cb ? ca > cb : vkey[a] < vkey[b];
            }
            return pcost[a] > pcost[b];
        });
        std::vector<int2> pitems;
        std::vector<size_t> first_item(pjobs.size(), 0);
        groups.clear();
        for (int j : porder) {
            if (groups.empty() || groups.back().variant != vkey[j])
                groups.push_back(ProjGroup{vkey[j], (int)pitems.size(), 0, 0});
            ProjGroup& G = groups.back();
            const ProjJob& J = pjobs[j];
            G.lds = std::max(G.lds, tile_lds_bytes(J.g.window_bins, tile_bin_bytes((J.flags & kJobPhase) != 0, J.g.fixed),
                                                   J.g.tile_lines, J.g.list_cap, J.g.list_packed));
            first_item[j] = pitems.size();
            const int n = J.blocks_per_slice * J.g.slices;
            for (int i = 0; i < n; i++) pitems.push_back(make_int2(j, i));
            G.count = (int)pitems.size() - G.first;
        }
        // convolution items by kernel variant, longest transforms first within one; a
        // launch per variant, its LDS that of its largest job
        std::vector<int> cvar(cjobs.size());
        for (size_t j = 0; j < cjobs.size(); j++)
            cvar[j] = conv_variant(cjobs[j], conv_block_threads(cjobs[j].fft_len));
        std::vector<int> corder(cjobs.size());
        std::iota(corder.begin(), corder.end(), 0);
        std::stable_sort(corder.begin(), corder.end(), [&](int a, int b) {
            if (cvar[a] != cvar[b]) return cvar[a] > cvar[b];
            return cjobs[a].fft_len > cjobs[b].fft_len;
        });
        std::vector<int2> citems;
        cgroups.clear();
        for (int j : corder) {
            if (cgroups.empty() || cgroups.back().variant != cvar[j])
                cgroups.push_back(ConvGroup{cvar[j], (int)citems.size(), 0, 0});
            ConvGroup& C = cgroups.back();
            C.lds = std::max(C.lds, conv_lds_bytes(cjobs[j]));
            for (int l = 0; l < cjobs[j].num_lines; l++) citems.push_back(make_int2(j, l));
            C.count = (int)citems.size() - C.first;
        }

        grow(d_proj_jobs, pjobs.size() * sizeof(ProjJob));
        grow(d_proj_items, pitems.size() * sizeof(int2));
        grow(d_fill, fills.size() * sizeof(FillSeg));
        grow(d_order_segs, ordered.size() * sizeof(OrderSeg));
        grow(d_conv_items, citems.size() * sizeof(int2));
        grow(d_conv_jobs, cjobs.size() * sizeof(ConvParams));
        grow(d_counters, std::max<size_t>(1, groups.size()) * sizeof(int));  // one per projection group
        grow(d_noise, cjobs.size() * sizeof(uint32_t));
        for (size_t k = 0; k < cjobs.size(); k++) cjobs[k].noise_counter_ptr = d_noise.as<uint32_t>() + k;
        std::vector<OrderSeg> order_segs;
        for (const auto& o : ordered) {
            const ProjJob& J = pjobs[o.first];
            order_segs.push_back(OrderSeg{o.second, d_proj_items.as<int2>() + first_item[o.first], J.blocks_per_slice,
                                          J.g.slices});
        }
        Stager st(*this, cjobs.size() * sizeof(ConvParams));
        st.add(d_proj_jobs, pjobs.data(), pjobs.size() * sizeof(ProjJob));
        st.add(d_proj_items, pitems.data(), pitems.size() * sizeof(int2));
        st.add(d_fill, fills.data(), fills.size() * sizeof(FillSeg));
        st.add(d_order_segs, order_segs.data(), order_segs.size() * sizeof(OrderSeg));
        st.add(d_conv_items, citems.data(), citems.size() * sizeof(int2));
        st.add(d_noise, noise_init.data(), noise_init.size() * sizeof(uint32_t));
        st.upload();
        conv_host = cjobs;
        n_proj_items = (int)pitems.size();
        n_conv_items = (int)citems.size();
        n_fill = (int)fills.size();
        n_order_segs = (int)order_segs.size();
        planned = true;
        just_built = true;
    }
    std::vector<ConvParams> conv_host;

    // Point each job's IQ at its slice of `out` (device memory).  Waits for frames
    // in flight first: they read the job table being rewritten.
    float2* table_out = nullptr;
    void set_output(float2* out) {
        if (table_out == out) return;
        // frames in flight read the table, and an earlier upload may still read the
        // staging head; right after build() neither holds (it drained the stream and
        // its own uploads use the rest of the staging buffer)
        if (!just_built) hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        just_built = false;
        // the head of the staging buffer (build() keeps it for this table)
        const size_t bytes = conv_host.size() * sizeof(ConvParams);
        ensure_stage(bytes);
        ConvParams* cj = static_cast<ConvParams*>(h_stage);
        for (size_t i = 0; i < conv_host.size(); i++) {
            cj[i] = conv_host[i];
            cj[i].iq = out + plan[i].out_offset;
        }
        if (bytes) hip_check(hipMemcpyAsync(d_conv_jobs.p, cj, bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
        table_out = out;
    }

    void enqueue() {
        const int ng = (int)groups.size();
        hip_check(launch_fill_segments(d_fill.as<FillSeg>(), n_fill, d_counters.as<int>(), ng,
                                       d_noise.as<uint32_t>(), (int)plan.size(), stream),
                  "fill kernel");
        for (const Render& r : renders)
            hip_check(launch_render_spline(r.set, r.basis, r.out, stream), "render kernel");
        hip_check(launch_order_segments(d_order_segs.as<OrderSeg>(), n_order_segs, stream), "block order kernel");
        for (int g = 0; g < ng; g++) {
            const ProjGroup& G = groups[g];
            hip_check(launch_project_batch(G.variant, d_proj_jobs.as<ProjJob>(), d_proj_items.as<int2>() + G.first,
                                           G.count, d_counters.as<int>() + g, G.lds, workgroups, stream),
                      "batched projection kernel");
        }
        for (size_t g = 0; g < cgroups.size(); g++) {
            const ConvGroup& C = cgroups[g];
            hip_check(launch_convolve_batch(C.variant, d_conv_jobs.as<ConvParams>(), d_conv_items.as<int2>() + C.first,
                                            C.count, C.lds, stream),
                      "batched convolution kernel");
        }
    }

    void simulate_device(void* d_out, size_t capacity) {
        if (stale()) build();
        if (capacity < total_out) fail(BCSIM_ERR_CAPACITY, "output capacity too small");
        hipEvent_t a = nullptr, b = nullptr;
        if (store_details) {
            hip_check(hipEventCreate(&a), "hipEventCreate");
            hip_check(hipEventCreate(&b), "hipEventCreate");
            hip_check(hipEventRecord(a, stream), "hipEventRecord");
        }
        if (use_graph) {
            if (!exec || graph_out != d_out) {
                drop_graph();
                set_output(static_cast<float2*>(d_out));
                hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                try {
                    enqueue();
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
                hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
                graph_out = d_out;
            }
            hip_check(hipGraphLaunch(exec, stream), "hipGraphLaunch");
        } else {
            set_output(static_cast<float2*>(d_out));
            enqueue();
        }
        for (bcsim_sim* s : sims) s->noise_counter++;  // this frame's counter is in flight on the device
        if (store_details) {
            hip_check(hipEventRecord(b, stream), "hipEventRecord");
            hip_check(hipEventSynchronize(b), "hipEventSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
            debug["batch_frame_ms"].push_back(ms);
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    }
};
