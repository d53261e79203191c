// kc_internal.h -- shared declarations of the kcgpu library (not part of the ABI).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/kc.h"

namespace kc {

// Per-thread error message behind kc_last_error().
void set_error(const std::string &msg);
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Simple undirected graph as a host CSR (rows sorted, no duplicates, no loops,
// every edge stored in both directions).
struct HostCSR {
    uint32_t n = 0;
    std::vector<uint32_t> rowptr;  // n + 1
    std::vector<uint32_t> col;     // rowptr[n]
};

// Build the simple undirected CSR from (src, dst) pairs; n == 0 derives
// max id + 1.  Returns KC_OK or an error code.
int build_csr(uint32_t n, size_t ne, const uint32_t *src, const uint32_t *dst, HostCSR &out);

uint64_t fnv1a_csr(const HostCSR &g);

// Loaders (kc_io.cpp).
int load_file(const char *path, HostCSR &out);
int read_edges_reference(const char *path, std::vector<uint32_t> &src, std::vector<uint32_t> &dst);

// Measured plan (kc_count.hip): both orientations, each granularity counted on
// a sample of its tasks at this k; the fastest wins.
int autotune_measured(const kc_graph *g, uint32_t k, uint32_t degree_max_out, kc_plan *plan);

// Generators (kc_gen.cpp): edge lists (u, v) with u != v.
int generate(const kc_gen_params &p, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
             uint32_t &n);

}  // namespace kc

struct kc_graph {
    kc::HostCSR csr;
    uint32_t max_degree = 0;
};
