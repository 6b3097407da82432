// Host-side utilities shared by the graph, reordering and BVSS code.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "blest.h"

namespace blest {

// Error carrying a C-ABI status code; thrown inside the library and converted to a
// blest_status at the extern "C" boundary.
struct Error : std::runtime_error {
    blest_status code;
    Error(blest_status c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

constexpr uint32_t kNoVertex = 0xFFFFFFFFu;  // reference UNSIGNED_MAX (Common.cuh:44)

// ---------------------------------------------------------------- random numbers
// splitmix64 finaliser (Steele, Lea, Flood 2014).  The generators draw the i-th
// value of stream s as the i-th splitmix64 output from a state keyed on (seed, s),
// so every draw is a pure function of (seed, stream, index): results do not depend
// on thread count or scheduling.
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed + kGolden) ^ mix64(stream * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull);
}
inline uint64_t crand(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * kGolden); }
inline double to_unit(uint64_t x) { return double(x >> 11) * (1.0 / 9007199254740992.0); }

// ---------------------------------------------------------------- threading
inline unsigned hw_threads(unsigned requested = 0) {
    if (requested) return requested;
    unsigned t = std::thread::hardware_concurrency();
    return t ? t : 1;
}

// Static block partition of [0, n) over `threads` std::threads.  fn(begin, end, tid).
template <class F>
void parallel_blocks(uint64_t n, unsigned threads, F&& fn) {
    threads = std::max(1u, std::min<unsigned>(threads, unsigned(std::max<uint64_t>(1, n))));
    if (threads == 1 || n < 2) {
        fn(uint64_t(0), n, 0u);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(threads);
    std::exception_ptr err;
    std::atomic<bool> failed{false};
    for (unsigned t = 0; t < threads; ++t) {
        uint64_t b = n * t / threads, e = n * (t + 1) / threads;
        pool.emplace_back([&, b, e, t] {
            try {
                fn(b, e, t);
            } catch (...) {
                if (!failed.exchange(true)) err = std::current_exception();
            }
        });
    }
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

// Dynamic chunked loop over [0, n): fn(i, tid) for each i.
template <class F>
void parallel_dynamic(uint64_t n, unsigned threads, uint64_t chunk, F&& fn) {
    threads = std::max(1u, std::min<unsigned>(threads, unsigned(std::max<uint64_t>(1, n))));
    std::atomic<uint64_t> next{0};
    auto worker = [&](unsigned tid) {
        for (;;) {
            uint64_t b = next.fetch_add(chunk);
            if (b >= n) break;
            uint64_t e = std::min(n, b + chunk);
            for (uint64_t i = b; i < e; ++i) fn(i, tid);
        }
    };
    if (threads == 1) {
        worker(0);
        return;
    }
    std::vector<std::thread> pool;
    std::exception_ptr err;
    std::atomic<bool> failed{false};
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            try {
                worker(t);
            } catch (...) {
                if (!failed.exchange(true)) err = std::current_exception();
                next.store(n);
            }
        });
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

}  // namespace blest
