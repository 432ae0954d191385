// Host-side copies of frame outputs: a memcpy split over threads for large
// copies, and the test for pageable (not pinned, not device) host memory.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace bcsim_hip {

// memcpy split over a few threads for large copies (page faults of a fresh
// destination are taken in parallel too).
inline void parallel_copy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kChunk = (size_t)4 << 20;
    const int parts = (int)std::min<size_t>(8, bytes / kChunk);
    if (parts <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t step = (bytes + parts - 1) / parts;
    std::vector<std::thread> th;
    for (int i = 1; i < parts; i++) {
        const size_t off = (size_t)i * step;
        const size_t n = std::min(step, bytes - std::min(bytes, off));
        th.emplace_back([=] { std::memcpy((char*)dst + off, (const char*)src + off, n); });
    }
    std::memcpy(dst, src, std::min(step, bytes));
    for (auto& t : th) t.join();
}

// Is p ordinary pageable host memory (neither pinned nor device memory)?
inline bool host_pageable(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // unknown to the runtime: pageable
        return true;
    }
    return a.type == hipMemoryTypeUnregistered;
}

}  // namespace bcsim_hip
