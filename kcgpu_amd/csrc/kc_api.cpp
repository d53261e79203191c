// kc_api.cpp -- host half of the C ABI: graph handles, loaders, generators,
// the -q spec parser.
#include <cstdlib>
#include <cstring>
#include <new>

#include "kc_internal.h"

using namespace kc;

static int finish_graph(HostCSR &&csr, kc_graph **out)
{
    kc_graph *g = new (std::nothrow) kc_graph;
    if (!g) return fail(KC_ERR_NOMEM, "graph handle");
    g->csr = std::move(csr);
    uint32_t md = 0;
    for (uint32_t v = 0; v < g->csr.n; ++v) {
        uint32_t d = g->csr.rowptr[v + 1] - g->csr.rowptr[v];
        if (d > md) md = d;
    }
    g->max_degree = md;
    *out = g;
    return KC_OK;
}

extern "C" {

int kc_load(const char *path, unsigned flags, kc_graph **out)
{
    (void)flags;
    if (!path || !out) return fail(KC_ERR_ARG, "kc_load: null argument");
    *out = nullptr;
    HostCSR csr;
    int rc = load_file(path, csr);
    if (rc) return rc;
    return finish_graph(std::move(csr), out);
}

int kc_read_edges_reference(const char *path, uint32_t **src, uint32_t **dst, size_t *count)
{
    if (!path || !src || !dst || !count) return fail(KC_ERR_ARG, "kc_read_edges_reference: null argument");
    std::vector<uint32_t> s, d;
    int rc = read_edges_reference(path, s, d);
    if (rc) return rc;
    size_t n = s.size();
    *src = (uint32_t *)std::malloc(sizeof(uint32_t) * (n ? n : 1));
    *dst = (uint32_t *)std::malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!*src || !*dst) {
        std::free(*src);
        std::free(*dst);
        return fail(KC_ERR_NOMEM, "edge buffers");
    }
    if (n) {
        std::memcpy(*src, s.data(), sizeof(uint32_t) * n);
        std::memcpy(*dst, d.data(), sizeof(uint32_t) * n);
    }
    *count = n;
    return KC_OK;
}

void kc_free_buffer(void *p) { std::free(p); }

int kc_graph_from_edges(uint32_t n, size_t nedges, const uint32_t *src, const uint32_t *dst, kc_graph **out)
{
    if (!out || (nedges && (!src || !dst))) return fail(KC_ERR_ARG, "kc_graph_from_edges: null argument");
    *out = nullptr;
    HostCSR csr;
    int rc = build_csr(n, nedges, src, dst, csr);
    if (rc) return rc;
    return finish_graph(std::move(csr), out);
}

int kc_graph_from_csr(uint32_t n, const uint32_t *rowptr, const uint32_t *col, kc_graph **out)
{
    if (!out || !rowptr || (!col && rowptr[n] != 0)) return fail(KC_ERR_ARG, "kc_graph_from_csr: null argument");
    *out = nullptr;
    if (rowptr[0] != 0) return fail(KC_ERR_ARG, "kc_graph_from_csr: rowptr[0] != 0");
    for (uint32_t v = 0; v < n; ++v)
        if (rowptr[v + 1] < rowptr[v]) return fail(KC_ERR_ARG, "kc_graph_from_csr: rowptr decreases at %u", v);
    size_t nnz = rowptr[n];
    std::vector<uint32_t> s(nnz);
    for (uint32_t v = 0; v < n; ++v)
        for (uint32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
            if (col[e] >= n) return fail(KC_ERR_ARG, "kc_graph_from_csr: column %u >= n", col[e]);
            s[e] = v;
        }
    HostCSR csr;
    int rc = build_csr(n, nnz, s.data(), col, csr);
    if (rc) return rc;
    return finish_graph(std::move(csr), out);
}

int kc_generate(const kc_gen_params *p, kc_graph **out)
{
    if (!p || !out) return fail(KC_ERR_ARG, "kc_generate: null argument");
    *out = nullptr;
    std::vector<uint32_t> s, d;
    uint32_t n = 0;
    int rc = generate(*p, s, d, n);
    if (rc) return rc;
    HostCSR csr;
    rc = build_csr(n, s.size(), s.data(), d.data(), csr);
    if (rc) return rc;
    return finish_graph(std::move(csr), out);
}

int kc_graph_get_info(const kc_graph *g, kc_graph_info *info)
{
    if (!g || !info) return fail(KC_ERR_ARG, "kc_graph_get_info: null argument");
    info->n = g->csr.n;
    info->nnz = g->csr.col.size();
    info->max_degree = g->max_degree;
    info->checksum = fnv1a_csr(g->csr);
    return KC_OK;
}

int kc_graph_get_csr(const kc_graph *g, uint32_t *rowptr, uint32_t *col)
{
    if (!g) return fail(KC_ERR_ARG, "kc_graph_get_csr: null graph");
    if (rowptr) std::memcpy(rowptr, g->csr.rowptr.data(), sizeof(uint32_t) * (g->csr.n + 1ull));
    if (col && !g->csr.col.empty()) std::memcpy(col, g->csr.col.data(), sizeof(uint32_t) * g->csr.col.size());
    return KC_OK;
}

void kc_graph_free(kc_graph *g) { delete g; }

// Rules measured on gfx950 (kc_count on generated graphs): pivoting with
// closed-form tails is never slower than the orientation DFS, one lane per
// candidate (P = 1) is the fastest mapping, a task whose vertex set fits one
// 64-bit word runs in the smallest register kernel.  So: keep the cheap degree
// orientation when it already bounds every out-degree by 64 (no peel needed),
// otherwise peel for the degeneracy order, which bounds out-degrees by the
// max core; split work per edge when out-degrees are heavy-tailed.
int kc_autotune(const kc_graph *g, uint32_t k, kc_plan *plan)
{
    if (!g || !plan) return fail(KC_ERR_ARG, "kc_autotune: null argument");
    const HostCSR &c = g->csr;
    uint32_t maxout = 0;
#pragma omp parallel for reduction(max : maxout) schedule(static, 4096)
    for (uint32_t u = 0; u < c.n; ++u) {
        const uint32_t du = c.rowptr[u + 1] - c.rowptr[u];
        uint32_t out = 0;
        for (uint32_t e = c.rowptr[u]; e < c.rowptr[u + 1]; ++e) {
            const uint32_t v = c.col[e], dv = c.rowptr[v + 1] - c.rowptr[v];
            out += (du < dv || (du == dv && u < v));
        }
        maxout = out > maxout ? out : maxout;
    }
    plan->degree_max_out = maxout;
    plan->algo = KC_ALGO_PIVOT;
    plan->part_size = 1;
    if (maxout <= 64 || k < 4) {
        // every task fits one word in the degree order: no peel, node tasks
        plan->orient_mode = maxout <= 64 ? KC_ORIENT_DEGREE : KC_ORIENT_DEGEN;
        plan->gran = maxout <= 64 ? KC_GRAN_NODE : KC_GRAN_EDGE;
        return KC_OK;
    }
    // larger tasks: which orientation and granularity is cheaper depends on
    // the graph (measured: degree order and edge tasks on a Chung-Lu graph,
    // degeneracy order and node tasks on R-MAT), so it is measured
    return kc::autotune_measured(g, k, maxout, plan);
}

// -q grammar (include/Config.h:235-271): [o|p]<P>[b|n|h]
int kc_parse_spec(const char *spec, kc_query *q)
{
    if (!spec || !q) return fail(KC_ERR_ARG, "kc_parse_spec: null argument");
    size_t len = std::strlen(spec);
    if (len < 2) return fail(KC_ERR_ARG, "bad -q spec '%s': expected [o|p]<P>[b|n]", spec);
    uint32_t algo;
    if (spec[0] == 'o') algo = KC_ALGO_ORIENT;
    else if (spec[0] == 'p') algo = KC_ALGO_PIVOT;
    else return fail(KC_ERR_ARG, "bad -q spec '%s': first letter must be o or p", spec);
    size_t i = 1;
    uint32_t ps = 0;
    while (i < len && spec[i] >= '0' && spec[i] <= '9' && ps < 1000) ps = ps * 10 + (uint32_t)(spec[i++] - '0');
    if (i == 1) return fail(KC_ERR_ARG, "bad -q spec '%s': missing partition size", spec);
    if (ps == 0 || ps > 64 || (ps & (ps - 1)))
        return fail(KC_ERR_ARG, "bad -q spec '%s': partition size must be 1,2,4,8,16,32 or 64", spec);
    uint32_t bin = 1;
    if (i < len) {
        if (spec[i] == 'n') bin = 0;
        else if (spec[i] == 'b' || spec[i] == 'h') bin = 1;
        else return fail(KC_ERR_ARG, "bad -q spec '%s': suffix must be b, n or h", spec);
        ++i;
    }
    if (i != len) return fail(KC_ERR_ARG, "bad -q spec '%s': trailing characters", spec);
    q->algo = algo;
    q->part_size = ps;
    q->binary = bin;
    return KC_OK;
}

}  // extern "C"
