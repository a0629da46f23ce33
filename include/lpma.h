/*
 * lpma.h -- C ABI of the MI355X-native batched edge-update store.
 *
 * A dynamic directed graph is kept as a packed memory array (PMA) of 64-bit
 * keys src<<32|dst in HBM: a gapped, sorted slot array with LPMA_KEY_NONE in
 * the gaps, one row-marker key (r<<32 | 0xFFFFFFFF) per vertex, and u32 row
 * offsets that point at the marker slots.  Every call that takes device
 * pointers expects memory on the handle's device; `stream` is a hipStream_t
 * (NULL = the default stream).  All calls return an lpma_status.
 *
 * Ownership: the caller owns every buffer it passes; the library never
 * writes to input buffers (the reference sorts and compacts the caller's
 * vectors in place, rpma.cuh:1557, :453).  The device view returned by
 * lpma_get_view is owned by the handle and stays valid until the next
 * mutating call on that handle.  One handle = one writer: calls on the same
 * handle must not overlap, and consecutive calls on it must be stream-ordered
 * (the same stream, or the previous call's stream synchronised first: a call
 * may return with its last kernels still queued).  Errors are returned, never
 * thrown across the ABI;
 * a failed update leaves the structure unchanged (every allocation of an
 * update precedes its first change), except after LPMA_ERR_DEVICE: a HIP
 * runtime failure leaves the handle undefined, and it must be destroyed.
 *
 * Each entry point names the reference interface it replaces
 * (file:line in pkumod/LPMA).
 */
#ifndef LPMA_H
#define LPMA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LPMA_KEY_NONE 0xFFFFFFFFFFFFFFFFull /* empty slot, rpma.cuh:38 / gpma.cuh:30 */
#define LPMA_MARKER_DST 0xFFFFFFFFu         /* dst of a row marker, gpma.cuh:34 (COL_IDX_NONE) */

typedef enum lpma_status {
    LPMA_OK = 0,
    LPMA_ERR_INVALID_ARGUMENT = 1, /* null handle/pointer, bad configuration */
    LPMA_ERR_INVALID_KEY = 2,      /* src or dst >= num_vertices (covers markers and KEY_NONE) */
    LPMA_ERR_OUT_OF_MEMORY = 3,    /* device allocation failed or the pool budget is exhausted */
    LPMA_ERR_DEVICE = 4,           /* a HIP runtime call failed */
    LPMA_ERR_TOO_LARGE = 5,        /* batch or capacity beyond the u32 slot/offset range */
    LPMA_ERR_NOT_BUILT = 6         /* the device code is not available */
} lpma_status;

typedef enum lpma_op { LPMA_OP_DELETE = 0, LPMA_OP_INSERT = 1 } lpma_op;

typedef struct lpma lpma_t;
typedef struct lpma_group lpma_group_t;
typedef struct lpma_snapshot lpma_snapshot_t;

/* Structure configuration.  The reference hard-codes 32-slot segments
 * (rpma.cuh:1769) and density bounds lower = 0.42*len, upper = len at every
 * level (rpma.cuh:76-79, :164-165); those are the defaults.  Unequal leaf and
 * root values interpolate linearly per level (the reference's commented-out
 * variant, rpma.cuh:144-163). */
typedef struct lpma_config {
    uint32_t segment_slots; /* power of two in [2, 1024]; default 32 */
    double density_lower_leaf;
    double density_lower_root;
    double density_upper_leaf;
    double density_upper_root;
    int device; /* HIP device ordinal; -1 = current device */
} lpma_config_t;

/* Device view of the gapped CSR (GPMA::keys / row_offset, gpma.cuh:41-62). */
typedef struct lpma_view {
    const uint64_t *keys;        /* capacity slots; LPMA_KEY_NONE in gaps */
    const uint32_t *row_offsets; /* num_vertices+1; [r+1] = slot of row r's marker, [0] = 0 */
    uint64_t capacity;           /* slots = segment_slots << tree_height */
    uint64_t live_keys;          /* edges + markers */
    uint64_t edges;
    uint32_t num_vertices;
    uint32_t segment_slots;
    uint32_t tree_height;        /* the array holds 2^tree_height segments */
    uint64_t version;            /* bumped by every successful update */
} lpma_view_t;

typedef struct lpma_stats {
    uint64_t updates_calls;
    uint64_t effective_updates;      /* after dedup and presence resolution */
    uint64_t windows_per_level[32];  /* windows rebalanced per level */
    uint64_t slots_per_level[32];    /* slots those windows cover */
    uint64_t rebuilds;               /* root rebalances and resizes */
    uint64_t rebuild_slots;
    uint64_t marker_moves;           /* row-offset writes */
    uint64_t alg_bytes_wave;         /* algorithmic bytes, 16W + 8k + 4m, per rebalance class */
    uint64_t alg_bytes_block;
    uint64_t alg_bytes_large;
    uint64_t alg_bytes_seg;
} lpma_stats_t;

/* Per-phase device time (ms) accumulated while profiling is on. */
typedef struct lpma_profile {
    double sort_ms;
    double locate_ms;     /* dedup + locate + compaction + density check */
    double elect_ms;
    double rebalance_wave_ms;
    double rebalance_block_ms;
    double large_ms;
    double rebuild_ms;
    double pyramid_ms;
    double rebalance_seg_ms;
    uint64_t launches_wave;
    uint64_t launches_block;
    uint64_t launches_seg;
    double graph_ms;          /* whole rounds replayed as one graph */
    uint64_t graph_launches;
    uint64_t graph_captures;
} lpma_profile_t;

/* Fills the reference defaults (32 slots, 0.42 / 1.0). */
void lpma_config_default(lpma_config_t *cfg);

/* init_csr_rpma (rpma.cuh:1807-1824) / init_rpma (rpma.cuh:1763-1795),
 * init_csr_gpma (gpma.cuh:844-856): empty graph on num_vertices vertices
 * (1 <= num_vertices < 2^31), one marker key per row.  cfg may be NULL. */
int lpma_create(lpma_t **out, uint32_t num_vertices, const lpma_config_t *cfg, void *stream);

/* destroy_rpma (rpma.cuh:1720-1732). */
int lpma_destroy(lpma_t *h);

/* update_rpma (rpma.cuh:1549-1718) / update_gpma (gpma.cuh:747-816):
 * apply one batch of n device keys (src<<32|dst) as inserts or deletes.
 * Duplicates in the batch collapse; inserting a present key and deleting an
 * absent key are no-ops (handle_del_mod, gpma.cuh:162-175). */
int lpma_update(lpma_t *h, const uint64_t *d_keys, uint64_t n, int op, void *stream);

/* One batch of inserts and deletes; deletes apply first, then inserts. */
int lpma_update_mixed(lpma_t *h, const uint64_t *d_ins, uint64_t n_ins, const uint64_t *d_del, uint64_t n_del,
                      void *stream);

/* `count` mixed batches applied in order as one pipeline run: the result
 * equals calling lpma_update_mixed(d_ins[i], n_ins[i], d_del[i], n_del[i])
 * for i = 0 .. count-1 (for each key the last batch touching it decides),
 * with one sort, one election and one rebalance pass for all of them.  The
 * reference applies one batch per update_rpma call (rpma.cuh:1549); this is
 * how a server merges queued write requests.  n_ins or n_del may be NULL. */
int lpma_update_sequence(lpma_t *h, const uint64_t *const *d_ins, const uint64_t *n_ins,
                         const uint64_t *const *d_del, const uint64_t *n_del, uint32_t count, void *stream);

/* Bulk load = one insert batch (significant_insert, gpma.cuh:720-744). */
int lpma_bulk_load(lpma_t *h, const uint64_t *d_keys, uint64_t n, void *stream);

/* Public members of GPMA / RPMA (gpma.cuh:41-62, rpma.cuh:60-96). */
int lpma_get_view(const lpma_t *h, lpma_view_t *out);

/* Copies the gapped view (capacity keys, num_vertices+1 row offsets) into
 * caller buffers; either pointer may be NULL. */
int lpma_copy_view(lpma_t *h, uint64_t *d_keys, uint32_t *d_row_offsets, void *stream);

/* Parity surface: the live edges (markers excluded) in key order, and
 * standard CSR row offsets (num_vertices+1; [v] = number of edges with
 * src < v).  Either pointer may be NULL.  d_edge_keys must hold
 * view.edges keys. */
int lpma_export_compacted(lpma_t *h, uint64_t *d_edge_keys, uint32_t *d_row_offsets, void *stream);

/* lpma_export_compacted without any allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Its scratch is a static
 * device array of about 8 MiB, held on each device the library runs on until
 * the process ends, so it handles structures of up to 2^21 segments
 * (LPMA_ERR_TOO_LARGE beyond).  Calls on one device share that array:
 * calls that overlap on one device race. */
int lpma_export_compacted_noalloc(lpma_t *h, uint64_t *d_edge_keys, uint32_t *d_row_offsets, void *stream);

/* get_data_rpma (rpma.cuh:1735-1761): all live keys including markers,
 * view.live_keys entries. */
int lpma_export_data(lpma_t *h, uint64_t *d_keys, void *stream);

/* to_csr_rpma (rpma.cuh:2249-2295): col_index holds dst, or src at each
 * row's marker position; row_offset[r+1] = position of row r's marker + 1
 * (so row_offset[v] = edges(src < v) + v).  col_index needs view.live_keys
 * entries, row_offset num_vertices+1. */
int lpma_to_csr(lpma_t *h, uint32_t *d_row_offset, uint32_t *d_col_index, void *stream);

/* edge_query_rpma (rpma.cuh:2167-2190): d_result[i] = 1 when d_keys[i] is a
 * live edge, else 0. */
int lpma_edge_query(lpma_t *h, const uint64_t *d_keys, uint64_t n, uint32_t *d_result, void *stream);

/* query_rpma (rpma.cuh:1974-2036) without the time-window filter: d_len[i] =
 * out-degree of d_srcs[i] (0 for ids >= num_vertices). */
int lpma_neighbor_count(lpma_t *h, const uint32_t *d_srcs, uint64_t n, uint32_t *d_len, void *stream);

/* Writes the neighbors (dst ids, ascending) of d_srcs[i] at
 * d_out[d_offsets[i] ...]; d_offsets typically the exclusive scan of
 * lpma_neighbor_count's lengths. */
int lpma_neighbor_gather(lpma_t *h, const uint32_t *d_srcs, uint64_t n, const uint32_t *d_offsets, uint32_t *d_out,
                         void *stream);

/* Breadth-first search from `src` over the gapped layout (the reference runs
 * nvGraph's BFS on a CSR copy, test_algos.cu:136-186).  d_levels (V entries)
 * receives the level of every vertex, 0xFFFFFFFF when unreached; `depth` the
 * number of levels, `edges_scanned` the live out-edges of the expanded
 * vertices.  Synchronises the stream once per level. */
int lpma_bfs(lpma_t *h, uint32_t src, uint32_t *d_levels, uint32_t *depth, uint64_t *edges_scanned, void *stream);

/* ---------------------------------------------------------------------------
 * Snapshots: a read-only device copy of the gapped view (keys, segment counts
 * and separators, row offsets) at one version.  Reads on a snapshot see that
 * version while updates go on through the handle, so long readers (BFS) run
 * beside writers.  The reference reads the live arrays only (query_rpma,
 * rpma.cuh:1974); there is no counterpart.  lpma_snapshot_create queues the
 * copies on `stream`; reads on any stream wait for them.  Destroy a
 * snapshot only after the reads on it have finished. */
uint64_t lpma_snapshot_bytes(const lpma_t *h);
int lpma_snapshot_create(lpma_snapshot_t **out, lpma_t *h, void *stream);
int lpma_snapshot_destroy(lpma_snapshot_t *s);
int lpma_snapshot_get_view(const lpma_snapshot_t *s, lpma_view_t *out);
int lpma_snapshot_neighbor_count(lpma_snapshot_t *s, const uint32_t *d_srcs, uint64_t n, uint32_t *d_len,
                                 void *stream);
int lpma_snapshot_neighbor_gather(lpma_snapshot_t *s, const uint32_t *d_srcs, uint64_t n, const uint32_t *d_offsets,
                                  uint32_t *d_out, void *stream);
int lpma_snapshot_edge_query(lpma_snapshot_t *s, const uint64_t *d_keys, uint64_t n, uint32_t *d_result,
                             void *stream);
int lpma_snapshot_bfs(lpma_snapshot_t *s, uint32_t src, uint32_t *d_levels, uint32_t *depth, uint64_t *edges_scanned,
                      void *stream);
/* Workgroups a search on this snapshot expands its frontier with (0: the
 * default, as many as the device holds at once).  Concurrent searches beside a
 * latency-sensitive writer use fewer, so the writer's kernels find free
 * compute units. */
int lpma_snapshot_set_search_width(lpma_snapshot_t *s, uint32_t workgroups);

/* PageRank over the gapped layout with nvGraph's formulation
 * (include/pagerank.hxx:66-88): from the uniform vector, until the L1 change
 * is below `tolerance` or `max_iter` iterations, then L1-normalised.  FP32
 * values and accumulation.  d_rank has V entries.  Synchronises the stream
 * once to size the per-call edge bins (V <= 2^27) and once per iteration.
 * Device scratch grows to about 8 B per edge plus 32 B per vertex. */
int lpma_pagerank(lpma_t *h, float damping, float tolerance, uint32_t max_iter, float *d_rank, uint32_t *iterations,
                  float *last_delta, void *stream);

/* Counts layout invariant violations on the device (0 = consistent). */
int lpma_check(lpma_t *h, uint64_t *violations, void *stream);

int lpma_get_stats(lpma_t *h, lpma_stats_t *out);
int lpma_reset_stats(lpma_t *h);

/* Per-phase HIP-event timing on the update stream (off by default). */
int lpma_set_profiling(lpma_t *h, int on);
int lpma_get_profile(lpma_t *h, lpma_profile_t *out);

/* Graph replay of update rounds (off by default for a handle): each round
 * is one captured hipGraph with grids sized for a batch capacity and a
 * single host synchronisation at its end, instead of the stream-ordered
 * kernel sequence that synchronises once after the window election.  The
 * graph is captured again when the batch capacity, a buffer or a member's
 * geometry changes.  Results are identical either way. */
int lpma_set_graph(lpma_t *h, int on);

/* Sizes the update buffers for batches of up to max_batch keys and, in graph
 * mode, captures the round graphs of every batch capacity up to it now, so
 * later rounds of that size capture nothing. */
int lpma_reserve(lpma_t *h, uint64_t max_batch, void *stream);

/* Makes root rebuilds and resizes up to max_edges live edges free of device
 * allocation.  A rebuild writes new arrays and keeps the old ones, so the
 * handle now allocates and holds one spare set of arrays of its current
 * height and two of every height above it up to the one max_edges needs
 * (about three times the arrays at that height); a shrink back to a height
 * held before allocates nothing either.  The scratch of large windows is
 * sized for max_edges.  The memory comes from the device pool and counts
 * against its budget; LPMA_ERR_OUT_OF_MEMORY leaves the structure
 * unchanged.  No counterpart in the reference, whose resize
 * allocates its new levels when it happens (rpma.cuh:1396-1415); for
 * callers whose latency bound covers resize batches. */
int lpma_reserve_capacity(lpma_t *h, uint64_t max_edges, void *stream);

/* Workgroup shape of the single-segment window rebalance: slots handled by
 * each lane (a power of two; 0 restores the default, 4 slots per lane up to
 * 256-slot segments, a full wave per segment above).  The lanes of one
 * window never exceed a wave and handle at most 16 slots each (larger
 * requests are lowered to 16); the LDS of a workgroup bounds it too.  No
 * counterpart in the reference (one warp per window, rpma.cuh:1226-1243). */
int lpma_set_launch(lpma_t *h, uint32_t seg_slots_per_lane);

/* Device pool shared by all handles on a device: cap on bytes held from
 * hipMalloc (0 = unlimited), and release of cached free blocks. */
int lpma_set_pool_budget(uint64_t bytes);
int lpma_pool_trim(void);
uint64_t lpma_pool_reserved_bytes(void);

/* Fault injection for failure-path tests (no counterpart in the reference,
 * whose allocations print and carry on, rpma.cuh:13-19): the nth allocation
 * from now (device pool blocks, cached or new, and pinned host buffers)
 * fails, and so does every later one until the next call, as when memory
 * runs out; calls that need them return LPMA_ERR_OUT_OF_MEMORY.  0 disables
 * it.  Process-wide.  When `remaining` is not NULL it receives the countdown
 * this call replaced (0: the earlier injection fired, or none was set). */
int lpma_inject_alloc_failure(uint64_t nth, uint64_t *remaining);

const char *lpma_strerror(int code);

/* ---------------------------------------------------------------------------
 * Served form: a request loop over one resident PMA (no counterpart in the
 * reference, whose drivers call update_rpma in a loop, test_update_rpma.cu).
 * Requests are submitted from any host thread with host buffers (copied at
 * submission) and served with continuous batching: all update requests
 * queued when the engine comes round are applied as one pipeline run in
 * submission order, all queued neighbor reads as one pass over the live
 * structure after them, and each BFS on a snapshot (the newest one, taken
 * at most snapshot_max_age_us earlier), on a worker stream beside the
 * engine.  The device memory
 * each request needs is charged against a byte budget; requests wait while
 * it has no room.  While a server runs, its handle must not be used
 * directly. */
typedef struct lpma_server lpma_server_t;

enum { LPMA_REQ_UPDATE = 0, LPMA_REQ_READ = 1, LPMA_REQ_BFS = 2 };

typedef struct {
    uint32_t bfs_workers;         /* searches running at once, default 2: with the engine's stream and
                                     the caller's they fill the 4 hardware queues HIP maps streams to */
    uint64_t hbm_budget_bytes;    /* 0: 90% of the device memory free at creation */
    uint64_t max_merge_keys;      /* keys of one merged update run, default 2^24 */
    uint32_t max_merge_requests;  /* update requests of one run, default 4096 */
    uint64_t max_merge_vertices;  /* vertices of one merged read pass, default 2^20 */
    uint64_t snapshot_max_age_us; /* a BFS reuses the newest snapshot while younger, default 50000;
                                     0: every BFS sees the version current at its admission */
    uint32_t search_workgroups;   /* workgroups of each search's frontier expansion, default 256 of
                                     the 2048 a lone search uses, so searches leave room for the writer */
} lpma_server_config_t;

typedef struct {
    int status;                   /* lpma_status of the request */
    int kind;                     /* LPMA_REQ_* */
    uint64_t applied_updates;     /* update requests applied when served (an update counts itself) */
    uint64_t version;             /* version of the structure it saw */
    double queue_us;              /* submission to start of service */
    double latency_us;            /* submission to completion */
    uint64_t count;               /* read: vertices; BFS: levels returned (0 or V) */
    const uint32_t *offsets;      /* read: count + 1 offsets into values */
    const uint32_t *values;       /* read: neighbor ids; BFS: levels (0xFFFFFFFF unreached) */
    uint32_t depth;               /* BFS levels */
    uint64_t edges_scanned;       /* BFS */
} lpma_server_result_t;

typedef struct {
    uint64_t submitted[3], completed[3];
    uint64_t update_runs, read_runs, snapshots, max_concurrent_bfs, admission_waits;
    uint64_t budget_bytes, in_use_bytes, peak_bytes, applied_updates;
    double update_us, read_us, dispatch_us;  /* engine time serving updates, reads, admitting searches */
    double stage_us, pipeline_us, finish_us; /* update service: staging copies, the update call, the final sync */
} lpma_server_stats_t;

void lpma_server_config_default(lpma_server_config_t *cfg);
int lpma_server_create(lpma_server_t **out, lpma_t *h, const lpma_server_config_t *cfg);
/* Serves every queued request, then stops. */
int lpma_server_destroy(lpma_server_t *s);
int lpma_server_submit_update(lpma_server_t *s, const uint64_t *h_ins, uint64_t n_ins, const uint64_t *h_del,
                              uint64_t n_del, uint64_t *ticket);
int lpma_server_submit_read(lpma_server_t *s, const uint32_t *h_vertices, uint64_t n, uint64_t *ticket);
int lpma_server_submit_bfs(lpma_server_t *s, uint32_t root, int want_levels, uint64_t *ticket);
/* Blocks until the request is done; the result's arrays stay valid until
 * lpma_server_release(ticket). */
int lpma_server_wait(lpma_server_t *s, uint64_t ticket, lpma_server_result_t *out);
int lpma_server_poll(lpma_server_t *s, uint64_t ticket, int *done, lpma_server_result_t *out);
int lpma_server_release(lpma_server_t *s, uint64_t ticket);
int lpma_server_get_stats(lpma_server_t *s, lpma_server_stats_t *out);

/* ---------------------------------------------------------------------------
 * Batched form: many independent PMAs updated by one pipeline run per round.
 * The reference updates one RPMA per update_rpma call (rpma.cuh:1549); it has
 * no counterpart for this.  A round sorts, locates, density-checks and elects
 * windows for every member at once (one host synchronisation), then runs
 * each rebalance class once for all members.  Each member ends exactly as if
 * it had been updated alone with lpma_update_mixed(d_ins[i], d_del[i]).
 *
 * Members must live on one device and outlive the group; a handle appears
 * at most once.  Members with different segment sizes, or too many ids to
 * fit above their sort keys, go to separate partitions (one run each).
 * An invalid key anywhere fails the whole round and changes no member. */
int lpma_group_create(lpma_group_t **out, lpma_t *const *members, uint32_t n);
int lpma_group_destroy(lpma_group_t *g);

/* One round: member i gets n_ins[i] inserts at d_ins[i] and n_del[i] deletes
 * at d_del[i] (deletes first).  Any of the four arrays may be NULL (no
 * updates of that kind); pointers with a zero count are ignored. */
int lpma_group_update(lpma_group_t *g, const uint64_t *const *d_ins, const uint64_t *n_ins,
                      const uint64_t *const *d_del, const uint64_t *n_del, void *stream);

/* Number of pipeline runs a round takes (partitions of the members). */
uint32_t lpma_group_partitions(const lpma_group_t *g);

/* Window statistics and phase timing of the group's rounds (the members'
 * own counters cover only their lpma_update calls). */
int lpma_group_get_stats(lpma_group_t *g, lpma_stats_t *out);
int lpma_group_set_profiling(lpma_group_t *g, int on);
/* Graph replay of the group's rounds (see lpma_set_graph); off by default. */
int lpma_group_set_graph(lpma_group_t *g, int on);
/* lpma_reserve for every partition of the group (max_batch = the largest
 * total batch of one round). */
int lpma_group_reserve(lpma_group_t *g, uint64_t max_batch, void *stream);
int lpma_group_set_launch(lpma_group_t *g, uint32_t seg_slots_per_lane);
int lpma_group_get_profile(lpma_group_t *g, lpma_profile_t *out);

/* ---------------------------------------------------------------------------
 * Input formats (host memory).  Generators use MT19937-64 with rejection
 * range reduction, so a seed fixes the output on every platform.
 */

/* `n` distinct uniform edges on V vertices, no self loops, in generation
 * order (rejection of repeats).  Returns 0 or an lpma_status. */
int lpma_gen_uniform(uint32_t V, uint64_t n, uint64_t seed, uint64_t *out);

/* `n` distinct R-MAT edges on V = 2^scale vertices with quadrant
 * probabilities (a, b, c, 1-a-b-c), no self loops, generation order. */
int lpma_gen_rmat(uint32_t scale, uint64_t n, double a, double b, double c, uint64_t seed, uint64_t *out);

/* R-MAT samples with replacement (repeats kept, self loops rejected): the
 * insert stream of the batched and served workloads. */
int lpma_gen_rmat_samples(uint32_t scale, uint64_t n, double a, double b, double c, uint64_t seed, uint64_t *out);

/* R-MAT samples generated on the current device into d_out (n keys, with
 * replacement, no self loops): sample i is a function of (seed, i) only, a
 * SplitMix64 counter stream (Steele, Lea & Flood 2014) whose 32-bit halves
 * pick the quadrant of one level each.  For the 2^28- and 2^29-edge
 * workloads, which the host generators would take minutes to draw. */
int lpma_gen_rmat_device(uint32_t scale, uint64_t n, double a, double b, double c, uint64_t seed, uint64_t *d_out,
                         void *stream);

/* Seeded Fisher-Yates shuffle of n keys in place. */
int lpma_shuffle(uint64_t *keys, uint64_t n, uint64_t seed);

/* n uniform edges on V vertices that are absent from the sorted array
 * `present` (np keys) and distinct from each other. */
int lpma_gen_absent(uint32_t V, const uint64_t *present, uint64_t np, uint64_t n, uint64_t seed, uint64_t *out);

/* Edgelist reader (README.md:7-17, test_update_rpma.cu:174-192): first line
 * "V E", then E lines "src dst".  With out == NULL only V and E are read.
 * max_edges limits how many edges are read (0 = all). */
int lpma_read_edgelist(const char *path, uint32_t *V, uint64_t *E, uint64_t max_edges, uint64_t *out);

/* Raw MT19937-64 stream (for known-answer checks). */
int lpma_mt64_fill(uint64_t seed, uint64_t n, uint64_t *out);

#ifdef __cplusplus
}
#endif

#endif /* LPMA_H */
