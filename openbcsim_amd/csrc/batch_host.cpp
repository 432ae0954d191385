// C-ABI of batched frames (bcsim_batch_*); the batch itself is in batch_state.h.
#include "batch_state.h"

extern "C" {

BCSIM_API bcsim_status bcsim_batch_create(bcsim_batch** out) {
    return guarded([&] {
        if (!out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null output handle");
        *out = new bcsim_batch();
    });
}

BCSIM_API bcsim_status bcsim_batch_destroy(bcsim_batch* b) {
    return guarded([&] { delete b; });
}

BCSIM_API bcsim_status bcsim_batch_set_parameter(bcsim_batch* b, const char* key, const char* value) {
    return guarded([&] {
        if (!b || !key || !value) fail(BCSIM_ERR_INVALID_ARGUMENT, "null argument");
        const std::string k = key;
        if (k == "use_graph") {
            b->use_graph = parse_bool(value);
            b->drop_graph();
        } else if (k == "store_kernel_details") {
            b->store_details = parse_bool(value);
        } else if (k == "autotune") {
            b->autotune = parse_bool(value);
            b->invalidate();
        } else if (k == "tile_lines") {
            const int v = parse_int(value);
            if (v < 0 || v > 32) fail(BCSIM_ERR_INVALID_ARGUMENT, "tile_lines must be in [0, 32]");
            b->tile_lines = v;
            b->invalidate();
        } else if (k == "items_per_workgroup") {
            const int v = parse_int(value);
            if (v < 1 || v > 1024) fail(BCSIM_ERR_INVALID_ARGUMENT, "items_per_workgroup must be in [1, 1024]");
            b->items_per_workgroup = v;
            b->invalidate();
        } else {
            fail(BCSIM_ERR_INVALID_ARGUMENT, "illegal parameter name: '" + k + "'");
        }
    });
}

BCSIM_API bcsim_status bcsim_batch_add(bcsim_batch* b, bcsim_sim* sim, int* job_index) {
    return guarded([&] {
        if (!b) fail(BCSIM_ERR_INVALID_ARGUMENT, "null batch handle");
        b->add(sim);
        if (job_index) *job_index = (int)b->sims.size() - 1;
    });
}

BCSIM_API bcsim_status bcsim_batch_clear(bcsim_batch* b) {
    return guarded([&] {
        if (!b) fail(BCSIM_ERR_INVALID_ARGUMENT, "null batch handle");
        b->sims.clear();
        b->device = -1;
        b->invalidate();
    });
}

BCSIM_API bcsim_status bcsim_batch_get_layout(bcsim_batch* b, size_t* total, int* num_jobs, size_t* offsets,
                                              int* lines, int* samples, int capacity) {
    return guarded([&] {
        if (!b) fail(BCSIM_ERR_INVALID_ARGUMENT, "null batch handle");
        if (b->stale()) b->build();
        if (total) *total = b->total_out;
        if (num_jobs) *num_jobs = (int)b->plan.size();
        for (int i = 0; i < (int)b->plan.size() && i < capacity; i++) {
            if (offsets) offsets[i] = b->plan[i].out_offset;
            if (lines) lines[i] = b->plan[i].lines;
            if (samples) samples[i] = b->plan[i].samples;
        }
    });
}

BCSIM_API bcsim_status bcsim_batch_simulate_device(bcsim_batch* b, void* d_out, size_t capacity) {
    return guarded([&] {
        if (!b || !d_out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null argument");
        b->simulate_device(d_out, capacity);
    });
}

BCSIM_API bcsim_status bcsim_batch_simulate(bcsim_batch* b, float* out, size_t capacity) {
    return guarded([&] {
        if (!b || !out) fail(BCSIM_ERR_INVALID_ARGUMENT, "null argument");
        if (b->stale()) b->build();
        if (capacity < b->total_out) fail(BCSIM_ERR_CAPACITY, "output capacity too small");
        DevBuf iq;
        iq.ensure(std::max<size_t>(b->total_out, 1) * sizeof(float2));
        const bool graph = b->use_graph;
        b->use_graph = false;  // a temporary output buffer is not worth a capture
        try {
            b->simulate_device(iq.p, b->total_out);
        } catch (...) {
            b->use_graph = graph;
            throw;
        }
        b->use_graph = graph;
        hip_check(hipMemcpyAsync(out, iq.p, b->total_out * sizeof(float2), hipMemcpyDeviceToHost, b->stream),
                  "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    });
}

BCSIM_API bcsim_status bcsim_batch_synchronize(bcsim_batch* b) {
    return guarded([&] {
        if (!b) fail(BCSIM_ERR_INVALID_ARGUMENT, "null batch handle");
        if (b->stream) hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    });
}

BCSIM_API bcsim_status bcsim_batch_get_debug_data(bcsim_batch* b, const char* key, double* out, size_t capacity,
                                                  size_t* count) {
    return guarded([&] {
        if (!b || !key) fail(BCSIM_ERR_INVALID_ARGUMENT, "null argument");
        auto it = b->debug.find(key);
        if (it == b->debug.end()) fail(BCSIM_ERR_INVALID_ARGUMENT, std::string("invalid debug data key: ") + key);
        if (count) *count = it->second.size();
        const size_t n = std::min(capacity, it->second.size());
        if (out && n) std::memcpy(out, it->second.data(), n * sizeof(double));
    });
}

}  // extern "C"
