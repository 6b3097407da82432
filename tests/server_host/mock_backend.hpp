// Driver hooks of the host-only server backend (mock_backend.cpp).
#pragma once

#include <cstddef>
#include <cstdint>

#include "blest.h"

constexpr uint32_t kFailRoot = 1000;  // a traversal from this root fails (in range on graph 0 only)

struct mock_report {
    uint64_t violations;  // contract breaches seen by the backend
    uint64_t live;        // queries not yet freed
    uint64_t peak_bytes;  // largest sum of live query footprints
    uint64_t created;     // queries created
    uint64_t calls;       // traversal calls
};

blest_device_graph* mock_graph_create(uint32_t n, uint32_t symmetric, uint32_t tag);
void mock_graph_free(blest_device_graph* g);
uint64_t mock_footprint(uint32_t max_sources);
void mock_set_budget(uint64_t bytes);
mock_report mock_report_get();
// "Device" result buffers: host memory registered with its size, so the backend can
// check that every result copy into one stays inside it (the library checks device
// destinations with hipMemGetAddressRange the same way).
void* mock_device_alloc(size_t bytes);
void mock_device_free(void* p);
