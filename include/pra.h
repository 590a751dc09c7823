/*
 * pra.h -- C ABI of the MI355X backend for parallel-RA's semi-naive hot path.
 *
 * The reference exposes no FFI: its "operator API" is the C++ class surface
 * that generated Datalog programs call (relation, RAM, parallel_join,
 * parallel_copy, parallel_acopy, parallel_copy_filter, LIE).  Each entry point
 * below replaces one piece of that surface; the C++ facade under
 * include/parallel_RA_inc.h forwards to it so such programs compile unchanged.
 *
 * Conventions: every function returns 0 on success or a negative PRA_E_* code;
 * pra_last_error() describes the most recent failure of a context.  Handles are
 * owned by the caller and freed with the matching *_destroy.  One context is
 * one LIE instance on one device and is not thread safe.  Rows are arrays of
 * (arity + 1) uint64 words, the last word being the intern id column.
 */
#ifndef PRA_H
#define PRA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pra_ctx pra_ctx;
typedef struct pra_rel pra_rel;
typedef struct pra_scc pra_scc;

/* relation versions: src/relation/balanced_hash_relation.h:13 (DELTA=0, FULL) */
#define PRA_DELTA 0
#define PRA_FULL 1
#define PRA_NEWT 2

/* copy_filter predicates over input columns 0 and 1 */
#define PRA_FILTER_EQ 0
#define PRA_FILTER_NEQ 1
#define PRA_FILTER_LT 2
#define PRA_FILTER_LE 3
#define PRA_FILTER_GT 4
#define PRA_FILTER_GE 5
#define PRA_FILTER_TRUE 6
#define PRA_FILTER_FALSE 7

#define PRA_OK 0
#define PRA_E_ARG -1        /* invalid argument or shape */
#define PRA_E_HIP -2        /* HIP runtime failure */
#define PRA_E_IO -3         /* file missing or malformed */
#define PRA_E_UNSUPPORTED -4
#define PRA_E_COLLISION -5  /* 64-bit key hash collision detected */
#define PRA_E_STATE -6      /* call out of order */
#define PRA_E_BUDGET -7     /* the context's device memory budget would be exceeded */
#define PRA_E_NOMEM -8      /* the device ran out of memory (after idle cached blocks were freed) */

/* Context = one LIE (src/lie/lie.h:13-153) on one device with `nbuckets`
 * virtual buckets, i.e. the layout of the reference run on `nbuckets` MPI
 * ranks (mpi_comm::create, src/comm/comm.cpp:22-28).  hbm_budget caps the
 * device bytes the context holds at any time (relation versions, tables and
 * scratch; 0 = no cap): a call that would need more fails with PRA_E_BUDGET
 * and leaves the context unusable except for pra_ctx_destroy. */
int pra_ctx_create(int device, int nbuckets, uint64_t hbm_budget, pra_ctx** out);
/* Several program instances in one context (no counterpart in the reference,
 * which runs one program per MPI job): pra_ctx_set_group(ctx, g) makes the
 * following pra_rel_create and pra_scc_create calls build instance g
 * (0 <= g < PRA_MAX_GROUPS; instance 0 before any call).  Each instance is a
 * separate LIE -- its own relations, SCCs, intern counters, dictionary and
 * join-pair count; rules and dependencies connect only its own relations and
 * SCCs.  With more than one instance, pra_execute runs every instance's LIE
 * in lock step: the current SCCs of all instances iterate together, each
 * phase of an iteration one batched launch over all of them, and each
 * instance's results equal those of its own run alone. */
#define PRA_MAX_GROUPS 65536
int pra_ctx_set_group(pra_ctx* ctx, int group);
/* join pairs of instance `group` so far */
int pra_ctx_group_joined(pra_ctx* ctx, int group, uint64_t* out);
/* One lock-step iteration instead of pra_execute's run to the end: every
 * instance without an SCC in progress starts its next runnable one, then one
 * iteration runs over all SCCs in progress.  *active = instances with work
 * left.  Instances built between steps join at the next step (the serving
 * loop admits requests this way). */
int pra_ctx_step(pra_ctx* ctx, uint64_t* active);
/* instance `group`: finished (no SCC in progress or runnable), its iterations
 * so far, and the device bytes its relations hold (any of the three may be
 * null) */
int pra_ctx_group_status(pra_ctx* ctx, int group, int* done, uint64_t* iterations, uint64_t* bytes);
/* free instance `group` (its relations and SCCs; their handles become
 * invalid) and reset its counters, so that the id can build a new program.
 * An instance with an SCC in progress is abandoned where it stands (between
 * steps; the other instances go on unaffected). */
int pra_ctx_release_group(pra_ctx* ctx, int group);
/* Between steps: give back the context's scratch (device buffers that hold
 * nothing from one iteration to the next, kept at the size of the largest
 * iteration so far); they grow again on demand.  A long-lived context that
 * admits instances under a memory budget calls it so that one large merged
 * iteration does not hold memory against later admissions. */
int pra_ctx_trim(pra_ctx* ctx);
/* Autotuner (no counterpart in the reference): variants of a step that give
 * identical results -- the dense claim table's size (2, 4 or 8 slots per
 * candidate of capacity) and the newt scatter's block width (256 or 1,024
 * threads) -- are timed on live single-output iterations, per device and band
 * of log2(work), three times each, and the cheapest is kept for the process.
 * PRA_AUTOTUNE=0 in the environment keeps the fixed defaults.  With
 * PRA_AUTOTUNE_FILE=<path>, the choices settled in earlier processes are read
 * from that file at the first context's creation (those cells run without
 * trials) and the file is rewritten whenever a cell settles.  The report has
 * one line per (device, knob, band): "device knob band choice mean-cost..."
 * (choice -1 while still trying; costs in ms per candidate).  *len receives
 * the report's length; up to cap bytes of it, NUL-terminated, go to buf. */
int pra_autotune_report(char* buf, uint64_t cap, uint64_t* len);
/* device bytes the context holds now and at most so far (no counterpart in
 * the reference; what budget admission measures) */
int pra_ctx_device_bytes(pra_ctx* ctx, uint64_t* live, uint64_t* peak);
int pra_ctx_destroy(pra_ctx* ctx);
const char* pra_last_error(pra_ctx* ctx);
/* LIE::set_batch_size (src/lie/lie.h:116) */
int pra_ctx_set_batch_size(pra_ctx* ctx, uint32_t batch_size);
/* the 20,000,000-output threshold of RAM::local_compute (src/RAM/RA_tasks.cpp:312) */
int pra_ctx_set_join_threshold(pra_ctx* ctx, int64_t threshold);
/* total (delta row, inner row) join pairs so far, counted before dedup */
int pra_ctx_joined_tuples(pra_ctx* ctx, uint64_t* out);
/* fixed-point iterations executed so far over all SCCs */
int pra_ctx_iterations(pra_ctx* ctx, uint64_t* out);
/* per-kernel timing with HIP events on the context stream: bit k of `mask`
 * times kernel class k (see pra_ctx_kernel_stat), 0 = off, -1 = all */
int pra_ctx_set_profiling(pra_ctx* ctx, int mask);
int pra_ctx_kernel_stat(pra_ctx* ctx, int idx, const char** name, uint64_t* launches, double* ms, double* bytes);
int pra_ctx_synchronize(pra_ctx* ctx);

/* relation(jcc, is_canonical, arity, tag, name, file, version)
 * (src/relation/balanced_hash_relation.h:90-102) */
int pra_rel_create(pra_ctx* ctx, uint32_t arity, uint32_t jcc, uint32_t tag, int canonical, pra_rel** out);
int pra_rel_destroy(pra_rel* rel);
/* populate_full / populate_delta from host rows
 * (src/relation/balanced_hash_relation.cpp:773-815) */
int pra_rel_load_rows(pra_rel* rel, const uint64_t* rows, uint64_t nrows, int version);
/* tsv_to_bin format: <path> raw LE rows, <path>.size "rows\ncols"
 * (src/IO/parallel_io.cpp:115-235, utility/tsv_to_bin.cpp) */
int pra_rel_load_bin(pra_rel* rel, const char* path, int version);
/* element count of a version (get_full_element_count and friends) */
int pra_rel_count(pra_rel* rel, int version, uint64_t* out);
/* rows of a version, bucket-major and lexicographic inside each bucket (the
 * trie walk order of src/relation/shmap_relation_exp.cpp:77-112) */
int pra_rel_download(pra_rel* rel, int version, uint64_t* rows, uint64_t cap, uint64_t* nrows);
/* write a version in the tsv_to_bin layout (<path> rows in download order,
 * <path>.size "rows\ncols"); with one bucket per rank this is the shared-file
 * dump of relation::parallel_IO (src/relation/balanced_hash_relation.cpp:109-235),
 * which concatenates the ranks' trie-ordered rows in rank order */
int pra_rel_write_bin(pra_rel* rel, const char* path, int version);
/* order-sensitive 64-bit checksum of a version's rows in download order:
 * sum over i of hash(row i, i) (bit-exact comparison of large results) */
int pra_rel_checksum(pra_rel* rel, int version, uint64_t* out);
/* device pointer to the sorted rows of a version (valid until the next call
 * that changes the relation) */
int pra_rel_device_rows(pra_rel* rel, int version, const uint64_t** rows, uint64_t* nrows);

/* RAM(bool run_to_fixed_point, int id) (src/RAM/RA_tasks.h:52) */
int pra_scc_create(pra_ctx* ctx, int fixpoint, int id, pra_scc** out);
int pra_scc_destroy(pra_scc* scc);
/* RAM::add_relation(relation*&, bool move_full_to_delta) (src/RAM/RA_tasks.cpp:34-39) */
int pra_scc_add_relation(pra_scc* scc, pra_rel* rel, int move_full_to_delta);
/* parallel_join(out, in0, ver0, in1, ver1, projection) (src/RA/parallel_join.h:55-58).  The join
 * matches on in0's join column count (src/RAM/RA_tasks.cpp:448); when in1's count differs, a pair is
 * found only where the reference's ranks meet it: in1's row in the same bucket under both counts. */
int pra_scc_add_join(pra_scc* scc, pra_rel* out, pra_rel* in0, int ver0, pra_rel* in1, int ver1, const int* proj,
                     int nproj);
/* parallel_copy(dest, src, ver, reorder) (src/RA/parallel_copy.h:28-32) */
int pra_scc_add_copy(pra_scc* scc, pra_rel* out, pra_rel* in, int ver, const int* reorder, int n);
/* parallel_acopy(dest, src, ver, reorder) (src/RA/parallel_acopy.h:45-49) */
int pra_scc_add_acopy(pra_scc* scc, pra_rel* out, pra_rel* in, int ver, const int* reorder, int n);
/* parallel_copy_filter(dest, src, ver, reorder, predicate) with the predicate
 * given as a PRA_FILTER_* code (src/RA/parallel_copy_filter.h:29-33) */
int pra_scc_add_copy_filter(pra_scc* scc, pra_rel* out, pra_rel* in, int ver, const int* reorder, int n, int op);
/* LIE::add_scc_dependance(before, after) (src/lie/lie.cpp:79-102) */
int pra_scc_add_dependency(pra_scc* before, pra_scc* after);
/* Run one SCC the way LIE::execute does (src/lie/lie.cpp:286-441); delta_hist
 * receives the global delta count after every batch (RAM::check_for_fixed_point,
 * src/RAM/RA_tasks.cpp:1072-1085). */
int pra_scc_run(pra_scc* scc, uint64_t* delta_hist, uint64_t cap, uint64_t* nbatches);
/* The same SCC run in steps, so a caller can act between batches (the
 * facade writes checkpoints there, src/lie/lie.cpp:353-423):
 * pra_scc_begin = the SCC prologue of LIE::execute (status-true relations
 * move full to delta unless move_full_to_delta is 0, as after a restart that
 * loaded the deltas, src/lie/lie.cpp:295-333); pra_scc_step = one batch
 * (RAM::execute_in_batches, src/RAM/RA_tasks.cpp:1121-1269; one iteration for
 * an SCC created with fixpoint 0), *delta = the global delta after it;
 * pra_scc_end = delta folded into full (RAM::insert_delta_in_full,
 * src/lie/lie.cpp:432). */
int pra_scc_begin(pra_scc* scc, int move_full_to_delta);
int pra_scc_step(pra_scc* scc, uint64_t* delta);
int pra_scc_end(pra_scc* scc);
/* number of batches run by the SCC and their delta counts */
int pra_scc_history(pra_scc* scc, uint64_t* out, uint64_t cap, uint64_t* n);

/* LIE::execute (src/lie/lie.cpp:197-448): run every SCC created on the
 * context in dependency order, first runnable in creation order. */
int pra_execute(pra_ctx* ctx);

/* Stable partition of device rows by bucket = tuple_hash(row[0:jcc]) % k
 * (k = the context's bucket count): the routing of input rows to their owner
 * rank (buffer_data_to_hash_buffer_col, src/IO/parallel_io.cpp:239-297) and of
 * the all-to-all tuple exchange (src/comm/all_to_allv_comm.cpp:64-142) as one
 * device step.  rows: n rows of w u64 words; out_rows: n rows, bucket 0's
 * first, each bucket's rows in input order; bucket_bounds: k + 1 u64 words,
 * bucket b = rows [bounds[b], bounds[b+1]).  All three are device pointers; the
 * work is ordered on the context's stream (pra_ctx_synchronize waits for it). */
int pra_partition_rows(pra_ctx* ctx, const uint64_t* rows, uint64_t n, uint32_t w, uint32_t jcc, uint64_t* out_rows,
                       uint64_t* bucket_bounds);
/* pra_partition_rows without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its scratch is a
 * static device array of 4,195,328 bytes that every GPU the library runs on
 * holds until the process ends; it takes at most 8,388,608 rows (PRA_E_ARG
 * beyond).  Same results as pra_partition_rows.  All calls on one GPU share
 * that GPU's scratch: calls that overlap on one GPU (from different contexts
 * or streams) race. */
int pra_partition_rows_noalloc(pra_ctx* ctx, const uint64_t* rows, uint64_t n, uint32_t w, uint32_t jcc,
                               uint64_t* out_rows, uint64_t* bucket_bounds);

/* Batched execution (no counterpart in the reference, which runs one program
 * per MPI job): pra_execute on n independent contexts of the same device,
 * taken in the given order by up to `concurrency` host worker threads; every
 * context has its own stream, so instances of uneven sizes overlap on the GPU.
 * rc[i] (optional) receives context i's status; returns the first failure. */
int pra_execute_many(pra_ctx** ctxs, int n, int concurrency, int* rc);

/* Serving (no counterpart in the reference): a request loop that solves
 * transitive-closure requests -- the program of tests/transitive_closure.cpp
 * on the given edge rows (m x 3 words: src, dst, row id) -- on one GPU with
 * continuous batching.  Up to `slots` requests are in flight as instances of
 * one context and advance together, one lock-step iteration at a time (each
 * phase one batched launch over all of them); between iterations finished
 * requests are retired and queued ones admitted, in arrival order, when the
 * requests in flight (each counted at the larger of its estimate and the bytes
 * its relations hold), the new request's estimate and the largest context
 * scratch seen after a step fit into `hbm_budget` (0 = no budget; an idle
 * server always admits).  est_bytes 0 = the server's
 * own estimate (random graphs with 1.3 edges per vertex).  Large requests
 * (estimate at least that of a 10,000-vertex graph; PRA_SERVE_LARGE_BYTES in
 * the environment at creation sets the bound, 0 turns the rule off) are in
 * flight one at a time, since each fills the device alone: one that waits
 * only for the large request in flight lets the requests behind it pass, one
 * that waits for memory holds the queue.  Under memory
 * pressure requests are restarted rather than failed: when the bytes held
 * pass the budget after an iteration, the context's scratch is given back
 * (pra_ctx_trim), then the most recently admitted requests in flight are
 * released and queued again at the head of the queue; when an
 * iteration runs the device out of memory (PRA_E_NOMEM / PRA_E_BUDGET) with
 * several requests in flight, all of them are queued again, the context is
 * rebuilt and the number in flight is halved, then raised by one per request
 * completed.  A request fails only when it fails alone.  keep_rows != 0 keeps
 * each result's rows on the host for pra_server_rows.  Times are seconds since
 * server creation; t_start is the last admission. */
typedef struct pra_server pra_server;
typedef struct {
    int status;
    uint64_t tuples, joined, iterations;
    double t_submit, t_start, t_done;
    uint32_t restarts;   /* times the request was released and queued again */
} pra_request_result;
int pra_server_create(int device, int nbuckets, int slots, uint64_t hbm_budget, int keep_rows, pra_server** out);
int pra_server_submit_tc(pra_server* srv, const uint64_t* edge_rows, uint64_t m, uint64_t est_bytes, uint64_t* ticket);
int pra_server_wait(pra_server* srv, uint64_t ticket, pra_request_result* out);
int pra_server_rows(pra_server* srv, uint64_t ticket, uint64_t* rows, uint64_t cap, uint64_t* nrows);
/* drains the queue, joins the workers, frees the server */
int pra_server_destroy(pra_server* srv);

#ifdef __cplusplus
}
#endif

#endif
