/*
 * kc.h -- C ABI of the MI355X (gfx950) exact k-clique counter.
 *
 * The reference (KCGPU) has no library interface: its path is reached through
 * the CLI in src/main.cu and the C++ templates of graph::SingleGPU_Kclique
 * (kclique/kclique.cuh:1302-2643).  Each entry point below names the
 * reference code it replaces.  Plain C types only; every call returns an int
 * status (KC_OK = 0) and never exits the process; kc_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Handles are opaque.  A handle is not thread-safe; distinct handles are.
 * All calls are synchronous: when a call returns, its device work is done.
 */
#ifndef KC_H
#define KC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define KC_OK            0
#define KC_ERR_ARG       1  /* invalid argument (bad k, bad -q spec, null) */
#define KC_ERR_IO        2  /* file missing or unreadable */
#define KC_ERR_FORMAT    3  /* malformed file, unknown extension */
#define KC_ERR_DEVICE    4  /* HIP runtime error, no usable gfx950 device */
#define KC_ERR_NOMEM     5  /* host or device allocation failed */
#define KC_ERR_OVERFLOW  6  /* the count is >= 2^64; result.overflow is set */
#define KC_ERR_BUDGET    7  /* admission refused: footprint above budget */
#define KC_ERR_NOTFOUND  8  /* unknown graph id / ticket */
#define KC_ERR_LIMIT     9  /* a fixed limit of the library was exceeded */

/* Message of the last failing call on this thread ("" if none). */
const char *kc_last_error(void);

/* ---- graphs ------------------------------------------------------------ */
typedef struct kc_graph kc_graph;        /* simple undirected graph (host CSR) */
typedef struct kc_oriented kc_oriented;  /* oriented CSR resident in HBM */

typedef struct {
    uint32_t n;           /* vertices (max id + 1) */
    uint64_t nnz;         /* CSR entries = 2 * undirected edges */
    uint32_t max_degree;
    uint64_t checksum;    /* FNV-1a 64 over rowptr[0..n] then col[0..nnz) as LE u32 */
} kc_graph_info;

/* kc_load flags */
#define KC_LOAD_SORT 1u   /* -s: accepted for drop-in; the builder always sorts */

/*
 * Load an edge list file and build the simple undirected CSR.
 * Replaces graph::EdgeListFile (include/FIleReader.cuh:33-476) plus
 * CSRCOO::from_edgelist (include/CSRCOO.cuh:37-80) and src/main.cu:73-131.
 * The extension selects the format as in the reference (.mtx, .tsv/.txt,
 * .bel).  Self-loops are dropped, every edge is stored in both directions and
 * duplicates are removed (the reference keeps duplicates and, for symmetric
 * MTX, only one direction of each edge).  Ids are kept as written (MTX stays
 * 1-based, so vertex 0 is isolated), n = max id + 1.
 */
int kc_load(const char *path, unsigned flags, kc_graph **out);

/*
 * Raw edge list exactly as the reference's read loop produces it
 * (src/main.cu:84-87 over EdgeListFile::get_edges with batches of 100),
 * including its defects: symmetric MTX keeps (col,row) and leaves a stale
 * slot per line, TSV emits both directions per line, BEL is not symmetrized.
 * For documentation tests of the reference's behaviour.  On success *src and
 * *dst are malloc'ed arrays of *count entries; release them with kc_free_buffer.
 */
int kc_read_edges_reference(const char *path, uint32_t **src, uint32_t **dst, size_t *count);
void kc_free_buffer(void *p);

/* ---- format converters (-m mtx-bel | tsv-bel | tsv-mtx | bel-mtx | txt-bel) */
#define KC_CONV_MTX_BEL 0  /* MTX -> BEL: symmetric entries mirrored, sorted, loops kept once */
#define KC_CONV_TSV_BEL 1  /* "dst src weight" triples -> BEL in file order */
#define KC_CONV_TSV_MTX 2  /* "dst src weight" triples -> general MTX "src dst weight" */
#define KC_CONV_BEL_MTX 3  /* BEL -> general MTX "src dst weight" */
#define KC_CONV_TXT_BEL 4  /* "a b" lines -> BEL of both directions, sorted, deduplicated, no loops */

/* kc_convert flags */
#define KC_CONV_REFERENCE_FLAGS 1u  /* mtx-bel: every '%' line re-reads the symmetric/pattern
                                       flags, as the reference does (a later comment clears them) */

/*
 * Convert between the on-disk formats.  Replaces graph::MtB_Writer
 * (include/FIleReader.cuh:480-933) as dispatched by src/main.cu:44-71
 * (-g source, -r destination).  Output bytes equal the reference's for
 * inputs on which the reference is well defined; where it is not (no size
 * line, an empty input, negative or >32-bit ids, an empty line in a .txt)
 * this returns KC_ERR_FORMAT, or skips the empty line.
 */
int kc_convert(int kind, const char *src_path, const char *dst_path, unsigned flags);

/* Build from an edge list (both arrays of length nedges; n = 0: max id + 1). */
int kc_graph_from_edges(uint32_t n, size_t nedges, const uint32_t *src, const uint32_t *dst,
                        kc_graph **out);

/*
 * Build from a CSR (rowptr[n+1], col[rowptr[n]]).  The library copies; the
 * caller keeps its arrays.  The CSR is cleaned like an edge list (symmetrized,
 * deduplicated, self-loops dropped).  Replaces the host half of
 * COOCSRGraph / to_csrcoo_device (include/GraphDataStructure.cuh:6-62).
 */
int kc_graph_from_csr(uint32_t n, const uint32_t *rowptr, const uint32_t *col, kc_graph **out);

int kc_graph_get_info(const kc_graph *g, kc_graph_info *info);
/* Copy the CSR out: rowptr needs n+1 entries, col needs nnz entries. */
int kc_graph_get_csr(const kc_graph *g, uint32_t *rowptr, uint32_t *col);
void kc_graph_free(kc_graph *g);

/* ---- seeded synthetic graphs ------------------------------------------- */
#define KC_GEN_ER        0  /* G(n, m): uniform endpoints, loops/duplicates rejected */
#define KC_GEN_CHUNG_LU  1  /* power law: w_i ~ (i+i0)^(-1/(gamma-1)), max expected degree = max_degree */
#define KC_GEN_RMAT      2  /* R-MAT scale/edge_factor/(a,b,c): 2^scale*edge_factor samples */
#define KC_GEN_BA        3  /* Barabasi-Albert: n vertices, attach edges per new vertex */

typedef struct {
    int kind;
    uint32_t n;            /* ER, CL, BA */
    uint64_t m;            /* ER, CL: unique undirected edges */
    double gamma;          /* CL degree exponent */
    double max_degree;     /* CL max expected degree */
    uint32_t scale;        /* RMAT */
    uint32_t edge_factor;  /* RMAT */
    double a, b, c;        /* RMAT quadrant probabilities (d = 1-a-b-c) */
    uint32_t attach;       /* BA */
    uint64_t seed;
} kc_gen_params;

/* Generators are deterministic in (params, seed) and independent of thread count. */
int kc_generate(const kc_gen_params *p, kc_graph **out);

/* ---- orientation ------------------------------------------------------- */
#define KC_ORIENT_DEGREE 0  /* -o degree: (deg u, u) < (deg v, v) */
#define KC_ORIENT_DEGEN  1  /* -o degen: (peel round u, u) < (peel round v, v) */

typedef struct {
    uint32_t n;
    uint64_t m;               /* oriented edges */
    uint32_t max_out_degree;
    uint32_t max_core;        /* degen only: the reference's "Max Core" */
    double preprocess_ms;     /* device time of upload-free orientation (HIP events) */
} kc_oriented_info;

/*
 * Orient on the device: degree key or k-core peel rounds, keep flags, per-row
 * kept counts, exclusive scan, stable per-row compaction.  Replaces
 * src/main.cu:164-235 with SingleGPU_Kcore::findKcoreIncremental_async
 * (kcore/kcore.cuh:442-584), init (include/main_support.cuh:54-128),
 * CUBSelect / CUBScanExclusive (include/cub_wrappers.cuh:11-179) and
 * warp_detect_deleted_edges (include/main_support.cuh:133-167).
 * The oriented CSR stays in HBM inside the returned handle.
 */
int kc_orient(const kc_graph *g, int mode, kc_oriented **out);

/*
 * Orient nb graphs, graph i by modes[i]; out[i] is what kc_orient(gs[i],
 * modes[i]) returns (same CSR, key and max out-degree).  Degree-order graphs
 * are preprocessed together: one upload of the concatenated CSRs and one
 * captured hipGraph of the orientation kernels over all of them, with the
 * oriented CSRs in one shared allocation (freed with the last of those
 * handles); preprocess_ms of each is the whole batch's device time.
 * Degeneracy-order graphs are oriented one by one.  On error no handle is
 * returned.  No reference counterpart (the reference reads one graph per
 * process, src/main.cu:73-235).
 */
int kc_orient_batch(const kc_graph *const *gs, const int *modes, size_t nb, kc_oriented **out);

/*
 * Build the simple undirected CSR of an edge list on the device (the same
 * graph as kc_graph_from_edges: self-loops dropped, both directions, rows
 * sorted, duplicates removed; n = 0: max id + 1) and orient it there, without
 * a host CSR: the edge list is uploaded once, then degree histogram, scan,
 * scatter, per-row sort and deduplication (a wave's bitonic sort in LDS, or a
 * bitmap over the vertex range for rows above 2048 entries), compaction and
 * orientation all run on the GPU.  Replaces CSRCOO::from_edgelist
 * (include/CSRCOO.cuh:37-80), the optional edge sort
 * (include/FIleReader.cuh:314-320) and src/main.cu:84-235.
 */
int kc_orient_edges(uint32_t n, size_t nedges, const uint32_t *src, const uint32_t *dst, int mode,
                    kc_oriented **out);

/* kc_graph_from_edges with the CSR built on the device (as in kc_orient_edges)
 * and copied back into a host graph. */
int kc_graph_from_edges_device(uint32_t n, size_t nedges, const uint32_t *src, const uint32_t *dst,
                               kc_graph **out);
int kc_oriented_get_info(const kc_oriented *o, kc_oriented_info *info);
/* Copy the oriented CSR (rowptr[n+1], col[m]) and the orientation key
 * (degree or peel round, key[n]) to the host; any pointer may be NULL. */
int kc_oriented_export(const kc_oriented *o, uint32_t *rowptr, uint32_t *col, uint32_t *key);
void kc_oriented_free(kc_oriented *o);

/*
 * Binary oriented-CSR cache: save an oriented graph and load it back onto
 * the device without loading, cleaning or orienting again.  The file is
 * little-endian: "KCOCSR1\0", u32 n, u32 mode, u64 m, u32 max_out_degree,
 * u32 max_core, u64 FNV-1a 64 of the three arrays, then rowptr[n+1], col[m],
 * key[n] (u32).  Loading checks the header, the checksum and that the CSR is
 * well formed (KC_ERR_FORMAT otherwise).  No reference counterpart: it cuts
 * src/main.cu:73-235 (read, build, orient) for repeated runs and uploads.
 */
int kc_oriented_save(const kc_oriented *o, const char *path);
int kc_oriented_load(const char *path, kc_oriented **out);

/* ---- counting ---------------------------------------------------------- */
#define KC_ALGO_ORIENT 0  /* -q o... */
#define KC_ALGO_PIVOT  1  /* -q p... */
#define KC_GRAN_NODE   0  /* -p node */
#define KC_GRAN_EDGE   1  /* -p edge */

typedef struct {
    uint32_t k;          /* clique size, 1..KC_MAX_K */
    uint32_t algo;       /* KC_ALGO_* */
    uint32_t gran;       /* KC_GRAN_* */
    uint32_t part_size;  /* lanes per partition: 1,2,4,8,16,32,64 */
    uint32_t binary;     /* 'b' = 1, 'n' = 0; both run the bitmap kernels */
    uint32_t auto_tune;  /* 1: choose algo/gran/part_size from the graph */
} kc_query;

#define KC_MAX_K 32

typedef struct {
    uint64_t count;       /* exact k-clique count (valid when overflow == 0) */
    uint32_t overflow;    /* 1 when the count does not fit in 64 bits */
    uint32_t algo, gran, part_size;  /* what ran (after auto-tune) */
    uint64_t tasks;       /* node or edge tasks that passed the filter */
    uint64_t alg_bytes;   /* algorithmic HBM bytes of those tasks */
    uint64_t tree_nodes;  /* DFS / SCT tree nodes visited */
    double count_ms;      /* device time of the counting kernels (HIP events) */
    double dominant_ms;   /* device time of the longest counting kernel */
    uint64_t dominant_bytes; /* algorithmic bytes of that kernel's tasks */
    uint32_t dominant_class; /* its size class (0: <=64, 1: <=128, 2: <=256, 3: <=512, 4: <=1024, 5: larger) */
    /* per size class: device time, algorithmic bytes, tasks, kernel variant
     * (0 none, 1 LDS bitmap, 2 register bitmap, 3 HBM bitmap, 4 LDS bitmap
     * shared by a block of waves, 5 HBM bitmap shared by a block of waves) */
    double class_ms[6];
    uint64_t class_bytes[6];
    uint64_t class_tasks[6];
    uint32_t class_kernel[6];
    double class_max_task_ms[6];  /* longest single task (wave wall clock) */
    double class_task_ms_sum[6];  /* sum of task durations over all waves */
} kc_result;

/*
 * Count k-cliques of the oriented graph.  Replaces the eight
 * SingleGPU_Kclique::findKclqueIncremental_{node,edge}_{binary_,pivot_,,nobin_pivot_}async<P>
 * drivers (kclique/kclique.cuh:1448-2623) and the dispatch of
 * src/main.cu:304-576.  Returns KC_ERR_OVERFLOW (with result->overflow = 1)
 * when the count is 2^64 or more.
 */
int kc_count(const kc_oriented *o, const kc_query *q, kc_result *result);

/*
 * kc_count over a task sample: only the tasks whose id is congruent to
 * offset modulo stride run (node tasks are numbered by vertex id, edge tasks
 * by oriented edge in CSR order; 0 <= offset < stride, stride 1 = kc_count).
 * The tasks a graph has are the reference's task list (node tasks:
 * kclique/kclique.cuh:1318-1361, edge tasks: :1365-1406); the sample lets a
 * CPU enumerator count the same tasks of a graph too large to count whole on
 * the host, so the two counts must be equal.  The result's count, tasks and
 * alg_bytes cover the sampled tasks only.
 */
int kc_count_sampled(const kc_oriented *o, const kc_query *q, uint32_t stride, uint32_t offset,
                     kc_result *result);

/*
 * Per-vertex counts (-m lkc): local[v] = number of k-cliques that contain v,
 * for every v < n (local has n entries).  result->count is the total count
 * (the sum of local[] is k times it); no entry can wrap unless the total
 * does, so result->overflow covers both.  Replaces
 * SingleGPU_Kclique_Local::findKclqueIncremental_{node,edge}_{binary,pivot}_async_local
 * (kclique/kclique_local.cuh:24-639, dispatched by src/main.cu:589-748);
 * unlike the reference it also takes 'n' specs (same kernels) and never
 * reads past local[n-1].
 */
int kc_count_local(const kc_oriented *o, const kc_query *q, uint64_t *local, kc_result *result);

/*
 * kc_count without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot call hipMalloc (a serving loop with a fixed
 * HBM footprint, a harness that forbids allocation during its run).  The
 * launch scratch (4 + 8 bytes per task, the binomial table, counters, and
 * HBM bitmap slices for tasks above 1024 vertices) is carved out of one
 * device array of KC_NOALLOC_SCRATCH_MIB MiB that the library's code object
 * holds on every device it is loaded on, until the process ends.  A query
 * whose scratch does not fit returns KC_ERR_LIMIT (for example edge tasks of
 * a graph above about 22 M oriented edges).  The count equals kc_count's;
 * the timing fields of the result stay 0 (no HIP events are created).  Calls
 * on one device share the array: they are serialized by a lock inside the
 * library, and the array keeps the last call's contents.
 */
#define KC_NOALLOC_SCRATCH_MIB 256
int kc_count_noalloc(const kc_oriented *o, const kc_query *q, kc_result *result);

/*
 * Parse the reference's -q grammar: [o|p]<P>[b|n|h] (include/Config.h:235-271).
 * Unlike the reference, a spec of two characters ("p8") is parsed too (the
 * reference silently ignores specs shorter than 3 characters) and an invalid
 * spec is an error instead of a silent default.  k, gran and auto_tune are
 * left untouched.
 */
int kc_parse_spec(const char *spec, kc_query *q);

/* ---- autotuner ----------------------------------------------------------- */
typedef struct {
    int orient_mode;        /* KC_ORIENT_* to pass to kc_orient */
    uint32_t algo, gran, part_size;
    uint32_t degree_max_out;  /* max out-degree the degree orientation would give */
} kc_plan;

/*
 * Pick orientation, granularity, algorithm and partition size for counting
 * k-cliques of g (the reference takes all four from the command line).  One
 * O(m) host pass over the CSR (include/Config.h:235-271 is what it replaces:
 * -o, -p and -q).  When the degree order keeps every out-degree at or below
 * 64, or k < 4, a rule decides (pivoting, P = 1, degree order and node tasks
 * when every task fits one 64-bit word).  Otherwise everything is measured
 * on the device: both orientations are built, a 1/64 sample of the tasks is
 * counted with pivoting for each (orientation, granularity) pair, and at
 * k <= 5 the same sample is also counted with the orientation DFS at
 * partition sizes 1, 8 and 64 on the winning pair; the fastest of all makes
 * the plan.  That needs a GPU and costs up to about one count.
 */
int kc_autotune(const kc_graph *g, uint32_t k, kc_plan *plan);

/* ---- batches: many (graph, query) pairs in one persistent launch ------- */
typedef struct {
    double total_ms;    /* device time: task sizing, binning and counting */
    double count_ms;    /* device time of the counting launches */
    uint64_t tasks;     /* tasks drained from the device queues */
    uint64_t queries;   /* queries that needed the GPU (k >= 3) */
} kc_batch_stats;

/*
 * Count nq queries, query i on graph gs[i] (graphs may repeat), in
 * persistent launches that drain one device queue of (query, vertex|edge)
 * tasks per (algorithm, size class) bin -- for small graphs a single launch.
 * results[i] gets exactly the count kc_count returns for query i alone.
 * All graphs must live on one device; the call uses the first graph's stream.
 * No counterpart in the reference (it counts one graph per process).
 */
int kc_batch_count(kc_oriented *const *gs, const kc_query *qs, size_t nq, kc_result *results,
                   kc_batch_stats *stats);

/* ---- serving ------------------------------------------------------------ */
typedef struct kc_server kc_server;

typedef struct {
    uint64_t budget_bytes;  /* HBM for resident CSRs + launch scratch; 0 = 90% of free memory */
    uint32_t workers;       /* batches in flight at once (one stream and 1/workers of the
                               device's block slots each); 0 = 4 */
    uint32_t max_batch;     /* queries per batch; 0 = 256 */
} kc_server_config;

typedef struct {
    uint64_t submitted, completed;  /* completed since the last kc_server_reset_stats */
    uint64_t batches;               /* worker batches + admissions into the resident launch */
    double mean_batch;              /* queries per batch or admission */
    double p50_ms, p99_ms, mean_ms; /* submit -> result latency */
    double queries_per_s;           /* completed / (last completion - first submit) */
    uint64_t graphs, resident_bytes, held_scratch_bytes;
    uint64_t resident_admitted;     /* queries counted by the resident launch (the rest by workers) */
    uint64_t joined_running;        /* of those, queries whose tasks were appended while it ran */
    uint64_t resident_launches;     /* times the resident launch was (re)started */
} kc_server_stats;

/*
 * Serving loop (no reference counterpart).  Graphs stay oriented in HBM;
 * kc_server_submit queues a query and returns.  Pivoting queries on graphs
 * whose tasks all fit 64 vertices are appended, as they arrive, to the task
 * queue of a resident persistent launch, whose waves drain them next to the
 * queries already running (continuous batching: a query admitted while
 * others are on the GPU joins the running launch when a query slot and
 * queue room are free).  Queries of at most 65,536 tasks go to an express
 * launch of 4 waves per CU, larger ones to a bulk launch, so a short query
 * is not held behind a long one.  Other queries are counted by worker
 * threads that batch every pending query fitting the memory budget into
 * persistent launches on their own streams.  A result is exactly what
 * kc_count returns for the query alone.
 */
int kc_server_create(const kc_server_config *cfg, kc_server **out);
/* Orient g and keep it resident; KC_ERR_BUDGET when it does not fit. */
int kc_server_add_graph(kc_server *s, const kc_graph *g, int orient_mode, uint32_t *gid);
/* Drop a graph: new queries on it fail with KC_ERR_NOTFOUND.  With nothing
 * in flight its memory is freed before the call returns (the free waits for
 * the device's running launches); otherwise a worker frees it after its
 * last query, and its bytes stay in the budget until then. */
int kc_server_remove_graph(kc_server *s, uint32_t gid);
/* Queue a query; returns at once with a ticket. */
int kc_server_submit(kc_server *s, uint32_t gid, const kc_query *q, uint64_t *ticket);
/* *done = 1 and *r filled when the ticket's result is ready (then the ticket is consumed). */
int kc_server_poll(kc_server *s, uint64_t ticket, kc_result *r, int *done);
/* Block until the ticket's result is ready. */
int kc_server_wait(kc_server *s, uint64_t ticket, kc_result *r);
/* Block until every submitted query has completed. */
int kc_server_drain(kc_server *s);
int kc_server_get_stats(kc_server *s, kc_server_stats *st);
int kc_server_reset_stats(kc_server *s);
void kc_server_destroy(kc_server *s);

/* ---- device ------------------------------------------------------------ */
typedef struct {
    char name[64];
    char arch[32];
    int cu_count;
    uint64_t hbm_bytes;
    uint64_t lds_per_block;
    int clock_khz;
} kc_device_info;

int kc_get_device_info(kc_device_info *info);

/* Select the HIP device (the reference's -d, include/Config.h:345) for
 * handles created after this call.  The device must be a gfx950. */
int kc_select_device(int ordinal);

/*
 * Release the library's device state: wait for its streams and destroy them
 * (handles carry no streams of their own).  Free every graph, oriented graph
 * and server first; handles created before this call must not be used after
 * it.  The next call that needs the device sets it up again.  Call it before
 * the process exits (the Python mirror does so from an atexit handler), so
 * that no HIP call is left for the runtime's teardown.  No reference
 * counterpart (its streams live until exit, include/utils_cuda.cuh:150-170).
 */
int kc_shutdown(void);

#ifdef __cplusplus
}
#endif
#endif /* KC_H */
