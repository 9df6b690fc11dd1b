/* gmk.h - C ABI of the MI355X maximal k-biplex enumerator (libgmk.so).
 *
 * The reference (GMkBPE) has no plugin or FFI layer: its boundary is the
 * `gmkbpe` process (CLI, .graph.bin input, result lines and time records) and
 * the internal seam enumerate_GPU / Generate_GPU.  Each entry point below names
 * the reference code it replaces.  All calls return GMK_OK or a negative
 * status; nothing calls exit().  A context owns one device and its buffers;
 * calls on one context must not overlap (different contexts may run in
 * parallel on different threads).
 */
#ifndef GMK_H
#define GMK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GMK_OK = 0,
  GMK_EINVAL = -1,     /* bad argument */
  GMK_EIO = -2,        /* file missing / malformed .graph.bin */
  GMK_EDEVICE = -3,    /* HIP runtime error (message in gmk_last_error) */
  GMK_ENOMEM = -4,     /* device allocation failed */
  GMK_ENOGRAPH = -5,   /* no graph loaded */
  GMK_EOVERFLOW = -6,  /* a fixed-capacity buffer overflowed (message says which) */
  GMK_EUNSUPPORTED = -7
};

/* Work order of the enumeration.
 *   GMK_MODE_PLAIN   one work item per launch sequence: one first vertex of
 *                    the thin enumerator, one seed partition (the reference's
 *                    one-candidate-at-a-time structure, src/main.cu:1078-1506)
 *   GMK_MODE_BATCHED work items fill each launch: thin tuples by the tens of
 *                    thousands, search states of all partitions in one device
 *                    stack per bitset width, expanded in batches that the
 *                    device plans from the last batch's outcome (one
 *                    workgroup per state, or a persistent grid taking states
 *                    from a device counter, as the tuning table says); items
 *                    in the degree order of the vertex IDs (the heuristic
 *                    order)
 *   GMK_MODE_LEARNED as BATCHED, with the one-vertex items ordered by the item
 *                    cost model set with gmk_set_item_model: best predicted
 *                    results per unit of device time first (single-graph
 *                    runs); batches order their graphs by the model set with
 *                    gmk_set_cost_model.  Without a model the heuristic order
 *                    is kept.  Every item is still enumerated exactly: the
 *                    order decides only which results a limited run reaches
 *                    first, and how the work packs into launches. */
enum { GMK_MODE_PLAIN = 0, GMK_MODE_BATCHED = 1, GMK_MODE_LEARNED = 2 };

/* Partition scheme.
 *   GMK_PART_AUTO    global partition when both sides fit one bitset tile,
 *                    else seeded
 *   GMK_PART_GLOBAL  the whole graph is one partition (small graphs)
 *   GMK_PART_SEEDED  per-seed 2-hop partitions for the results with both sides
 *                    >= 2k+1, plus the thin enumerators for the rest */
enum { GMK_PART_AUTO = 0, GMK_PART_GLOBAL = 1, GMK_PART_SEEDED = 2 };

typedef struct gmk_ctx gmk_ctx;

typedef struct {
  uint32_t k;            /* -k, default 1 (src/main.cu:1726) */
  uint32_t theta;        /* size lower bound on both sides, default 1 (src/main.cu:1585) */
  uint64_t limit;        /* -r, 0 = no limit (src/main.cu:1721) */
  int32_t mode;          /* GMK_MODE_* */
  int32_t partition;     /* GMK_PART_* */
  int32_t keep_results;  /* 1: keep the record stream for gmk_fetch_results */
  int32_t verbose;       /* -m */
  uint64_t result_capacity_words; /* 0 = automatic */
  int32_t dedup;         /* 1: insert every result into a device hash table (two
                            independent 64-bit hashes per result) and drop
                            repeats: a safety net, the decomposition emits none.
                            Limited runs size the table for the limit, unlimited
                            ones grow it as results arrive */
  uint32_t dedup_load_permille;   /* table load factor, 0 = the tuning table's */
} gmk_params;

typedef struct {
  uint64_t count;          /* results emitted (== limit when the limit was hit) */
  uint64_t fp_sum, fp_xor; /* order-free digest of the emitted set */
  uint64_t states;         /* search states expanded */
  uint64_t intersections;  /* bitset row intersections (AND + popcount) */
  uint64_t algo_bytes;     /* algorithmic bytes read + written by the kernels */
  uint64_t partitions;     /* partitions built */
  uint64_t launches;       /* kernel launches */
  uint64_t dedup_inserts, dedup_probes, dedup_dups;
  double ms;               /* enumeration window (after graph upload) */
  double ms_expand;        /* device time of all enumeration kernels */
  int32_t hit_limit;
  uint32_t spills;         /* times a state stack over 3/4 full moved its shallow half to the spill buffer */
  /* per kernel: 0 state expansion, 1 thin family construction,
   * 2 thin combination check, 3 partition construction */
  double ms_kernel[4];       /* device time (HIP events around each launch) */
  uint64_t launches_kernel[4];
  uint64_t bytes_kernel[4];  /* algorithmic bytes read + written */
  uint64_t dedup_undecided;  /* results kept because a just-claimed slot with the same
                                fingerprint had not yet shown its check word (dedup) */
  double ms_schedule;        /* GMK_MODE_LEARNED: host time of the item ordering (scoring
                                kernel, selection, sort), inside the window `ms` */
  double ms_reduce;          /* seeded runs: device time of the (lb-k)-core reduction and
                                the butterfly reduction before the seed partitions */
  uint64_t bytes_reduce;     /* its algorithmic bytes */
  uint64_t reduce_removed;   /* vertices the reductions removed */
  uint64_t widest_side;      /* most vertices on one side of any branch-and-bound partition
                                (up to 8192 per side a state's working set is in LDS, up
                                to 65536 in HBM) */
  uint64_t reduce_edges;     /* edges the butterfly reduction removed */
  double ms_build;           /* seeded runs: wall time building the seed partitions' vertex
                                sets (BPPivot constructSets, on the device) and placing them
                                in the partition tables, inside the window `ms` */
} gmk_stats;

/* Create a context on HIP device `device` (-d, src/main.cu:1765-1766). */
int gmk_create(int device, gmk_ctx** out);
void gmk_destroy(gmk_ctx* ctx);
const char* gmk_last_error(gmk_ctx* ctx);
/* Name of the context's device, NUL-terminated and cut to cap bytes (replaces
 * cudaGetDeviceProperties for the "Using GPU: <name>, GPU id: <d>" line,
 * src/main.cu:1768-1770). */
int gmk_device_name(gmk_ctx* ctx, char* out, uint32_t cap);

/* Load a .graph.bin file and upload its CSR (replaces load_graph_from_bin,
 * data/src/tmp_load.h:213-295, and the H2D copies at src/main.cu:1680-1690). */
int gmk_load_graph_bin(gmk_ctx* ctx, const char* path, uint32_t* n, uint32_t* n_left, uint64_t* nnz);

/* Upload a CSR held by the caller (copied; the caller keeps ownership).
 * off has n+1 entries; vertices [0, n_left) are left, [n_left, n) right;
 * neighbour lists must be sorted ascending and point to the other side. */
int gmk_set_graph_csr(gmk_ctx* ctx, uint32_t n, uint32_t n_left, const uint32_t* off, const uint32_t* adj);

/* Enumerate maximal k-biplexes (L, R) with |L|, |R| >= theta (replaces
 * enumerate_GPU, src/main.cu:999-1580, with Generate_GPU and its callees). */
int gmk_enumerate(gmk_ctx* ctx, const gmk_params* params, gmk_stats* stats);

/* Copy the record stream of the last gmk_enumerate run with keep_results = 1:
 * u32 words [|L|, L ascending, |R|, R ascending] per result, global IDs,
 * records in emission order.  *nwords receives the stream length; when
 * cap_words is smaller nothing is copied and GMK_EOVERFLOW is returned. */
int gmk_fetch_results(gmk_ctx* ctx, uint32_t* words, uint64_t cap_words, uint64_t* nwords);

/* Progress log of the last gmk_enumerate run: after each device launch, the
 * number of results emitted so far and the elapsed time in ms since the run
 * started.  Replaces the checkpoint records (count, ms) the reference takes at
 * result milestones (src/main.cu:1753-1761, include/for_generate.cuh:451-459):
 * a milestone n is reached at the first entry with count >= n.  *n receives
 * the log length; at most cap entries are copied. */
int gmk_fetch_progress(gmk_ctx* ctx, uint64_t* counts, double* ms, uint32_t cap, uint32_t* n);

/* Parse an edge list (`u v` per line) on `threads` host threads (0: all
 * cores), as graph_convert does (preprocess/graph_conver.cpp:505-552, same
 * rules and byte ranges).  *u and *v receive malloc'd arrays of *m raw IDs
 * each (free with gmk_free); the converter (gmkbe_amd/convert.py) relabels
 * them into the .graph.bin layout.  Host only; needs no device. */
int gmk_parse_edge_list(const char* path, int threads, uint32_t** u, uint32_t** v, uint64_t* m);
void gmk_free(void* p);

/* ---- Many graphs in one run -------------------------------------------- */

typedef struct {
  uint32_t n, n_left;      /* as gmk_set_graph_csr */
  const uint32_t* off;     /* n + 1 offsets (host memory) */
  const uint32_t* adj;     /* neighbour lists, ascending (host memory) */
  uint32_t k, theta;       /* per-graph k and size lower bound */
} gmk_graph_desc;

typedef struct {
  uint64_t count;          /* results kept for this graph (<= params->limit when set) */
  uint64_t fp_sum, fp_xor; /* order-free digest of the kept results */
  double ms;               /* GMK_MODE_PLAIN: device time of this graph's run */
  int32_t hit_cap;         /* 1 when the per-graph cap was reached */
  int32_t pad;
} gmk_graph_stats;

/* Enumerate many graphs in one call (the reference runs one graph per process).
 * Each graph with theta <= 2k is one global partition (both sides <= 65536),
 * each other graph runs as seed partitions.  params->limit is a
 * per-graph cap (0 = none); params->k / theta are ignored (per-graph values).
 *   GMK_MODE_PLAIN    graphs one after another, one block per state
 *   GMK_MODE_BATCHED  all graphs' states in one device stack per bitset width,
 *                     in batches planned on the device, graphs taken in
 *                     `order` (null = input order)
 *   GMK_MODE_LEARNED  as BATCHED, with the order given by the cost model set
 *                     with gmk_set_cost_model (most expensive first), or by
 *                     the size heuristic (edges x vertices) when none is set.
 * per_graph receives n_graphs entries. */
int gmk_enumerate_batch(gmk_ctx* ctx, const gmk_graph_desc* graphs, uint32_t n_graphs, const gmk_params* params,
                        const uint32_t* order, gmk_graph_stats* per_graph, gmk_stats* total);

/* ---- Learned cost model (no counterpart in the reference) ---------------
 * A small FP32 MLP, GMK_MLP_IN features -> 64 -> 64 -> 1 with ReLU, predicts
 * log(device ms) of a work item.  Training, inference and gradients run in
 * the library's own HIP kernels.  The model only orders work: results do not
 * depend on it. */
#define GMK_MLP_IN 8
#define GMK_MLP_HID 64
typedef struct {
  float w1[GMK_MLP_IN * GMK_MLP_HID];   /* [input][hidden] */
  float b1[GMK_MLP_HID];
  float w2[GMK_MLP_HID * GMK_MLP_HID];  /* [hidden][hidden] */
  float b2[GMK_MLP_HID];
  float w3[GMK_MLP_HID];
  float b3;
  float mu[GMK_MLP_IN], sd[GMK_MLP_IN]; /* input standardisation */
} gmk_mlp;

/* Features of a graph work item: log(1+nL), log(1+nR), log(1+|E|),
 * log(1+mean left degree), log(1+mean right degree), log(1+max degree),
 * log(1+paths of length two) (sum over vertices of C(deg, 2)), k. */
int gmk_graph_features(const gmk_graph_desc* g, float out[GMK_MLP_IN]);

/* Full-batch Adam on mean squared error; X is n x GMK_MLP_IN row-major (raw
 * features), y the targets.  Standardisation is fitted on X.  With n >= 20 a
 * seeded fifth of the samples is held out and the parameters with the lowest
 * held-out error (checked every 5 epochs) are returned (early stopping).
 * *loss receives the mean squared error over all n samples. */
int gmk_mlp_train(gmk_ctx* ctx, const float* X, const float* y, uint32_t n, uint32_t epochs, float lr, uint32_t seed,
                  gmk_mlp* model, float* loss);
int gmk_mlp_predict(gmk_ctx* ctx, const gmk_mlp* model, const float* X, uint32_t n, float* y);
/* Loss and gradient of the loss with respect to every parameter (mu/sd of the
 * gradient are left zero), for gradient checks. */
int gmk_mlp_grad(gmk_ctx* ctx, const gmk_mlp* model, const float* X, const float* y, uint32_t n, gmk_mlp* grad,
                 float* loss);
/* Cost model used by GMK_MODE_LEARNED batches (null: the size heuristic). */
int gmk_set_cost_model(gmk_ctx* ctx, const gmk_mlp* model);

/* ---- Learned schedule of single-graph runs (no counterpart in the reference)
 * A work item is a one-vertex tuple u of the thin enumerator: u, all of N(u)
 * and up to k vertices u misses, checked for maximality combination by
 * combination.  The item model predicts log(results per ms of device lane
 * time) of an item from its features.
 *
 * Features of items first .. first + count - 1 (global IDs, one side), for k:
 * log(1+d(u)), log(1+|other side| - d(u)), mean, max and standard deviation of
 * log(1+d(v)) over N(u), log(1+two-hop paths from u), log(1+min d(v)), k
 * (moments over at most 16 neighbours spread evenly over N(u)).  out is
 * count x GMK_MLP_IN. */
int gmk_item_features(gmk_ctx* ctx, uint32_t first, uint32_t count, uint32_t k, float* out);
/* Predictions of the item model set with gmk_set_item_model for items first ..
 * first + count - 1 (the scores the learned order sorts by; higher first). */
int gmk_item_predict(gmk_ctx* ctx, uint32_t first, uint32_t count, uint32_t k, float* out);
/* As gmk_item_predict, calling no allocator (no hipMalloc, hipFree or
 * allocating library routine): for callers that cannot allocate, such as
 * tight serving loops.  It has no counterpart in the reference.  Its device
 * scratch (the staged model and the scores) is a fixed array of the library,
 * one per GPU, held until the process ends: at most GMK_STATIC_ITEMS items per
 * call (larger counts return GMK_EINVAL).  Calls on one GPU share that array,
 * so calls that overlap on one GPU race; `out` is host memory. */
#define GMK_STATIC_ITEMS 65536
int gmk_item_predict_static(gmk_ctx* ctx, uint32_t first, uint32_t count, uint32_t k, float* out);
/* Measured yield and cost of items first .. first + count - 1 (one side of the
 * loaded graph; k = params->k, theta <= 1): the first `ranks` combinations of
 * each of an item's families are checked, and per item the maximal results
 * found and the device lane time spent (ms, summed over the lanes that checked
 * its combinations, from the kernels' wall clock) are returned.  This is the
 * method's own log that the item model trains on. */
int gmk_probe_items(gmk_ctx* ctx, const gmk_params* params, uint32_t first, uint32_t count, uint64_t ranks,
                    uint64_t* results, double* lane_ms);
/* Item model of GMK_MODE_LEARNED single-graph runs (null: heuristic order). */
int gmk_set_item_model(gmk_ctx* ctx, const gmk_mlp* model);

/* ---- Tuning (no counterpart in the reference) -----------------------------
 * Launch parameters of the branch-and-bound expansion per bitset width
 * (index l: 2^l words, l = 0..6), as chosen per size bucket by the autotuner
 * (gmkbe_amd/autotune.py, table under gmkbe_amd/tune/).  Results do not depend
 * on them.  Built-in defaults: one workgroup per state, 16384 states per
 * launch, dedup load 0.5, one lane per rank in the thin checks; the engine of
 * gmkbe_amd/engine.py loads the committed MI355X table instead. */
#define GMK_TUNING_VERSION 2
typedef struct {
  uint32_t version;                 /* GMK_TUNING_VERSION */
  uint32_t waves_per_cu[7];         /* batched modes: persistent waves per CU; 0 = one workgroup per state */
  uint32_t batch_states[7];         /* states per expansion launch, 64 .. 65536 */
  uint32_t dedup_load_permille;     /* dedup table load when gmk_params leaves it 0 (100 .. 900) */
  uint32_t check_form[4];           /* per k, thin checks: 0 one lane per rank, 1 waves taking 64-rank chunks */
} gmk_tuning;
int gmk_set_tuning(gmk_ctx* ctx, const gmk_tuning* tuning);
int gmk_get_tuning(gmk_ctx* ctx, gmk_tuning* tuning);

#ifdef __cplusplus
}
#endif
#endif /* GMK_H */
