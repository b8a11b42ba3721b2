// Glue between the C++ host API and the C ABI of the engine.
#pragma once
#include <genie/exception/exception.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "genie_hip.h"

namespace genie_amd {

// Map an engine status onto the reference's exception classes
// (src/genie/utility/cuda_macros.h:49-61 -> gpu_runtime_error; allocation
// failures -> gpu_bad_alloc, match.cu:461-463).
inline void check(int rc)
{
    if (rc == GENIE_OK)
        return;
    const std::string msg = genie_hip_last_error();
    switch (rc) {
    case GENIE_E_OOM: throw genie::exception::gpu_bad_alloc(msg);
    case GENIE_E_RUNTIME: throw genie::exception::gpu_runtime_error(msg);
    case GENIE_E_NOT_BUILT: throw genie::exception::cpu_runtime_error("table not built! " + msg);
    default: throw genie::exception::cpu_runtime_error(msg);
    }
}

// Run f(i) for i in [0, n) on up to 16 threads.
template <class F>
void parallel_for(size_t n, F f)
{
    const size_t T = std::min<size_t>(n, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    if (T <= 1) {
        for (size_t i = 0; i < n; ++i)
            f(i);
        return;
    }
    std::atomic<size_t> next(0);
    std::vector<std::thread> th;
    for (size_t t = 0; t < T; ++t)
        th.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;)
                f(i);
        });
    for (auto& x : th)
        x.join();
}

// Select the configured device once per process (utility/init.cu:7-34).
void ensure_device();
void set_device(int id);

}  // namespace genie_amd
