/*
 * mgb_groupby.h -- C ABI of the MI355X hash groupby-aggregate library
 * (libcudf_amd.so).
 *
 * Plain C: device pointers, sizes and status codes only. The C++ drop-in
 * (cudf::groupby::groupby, include/cudf/groupby.hpp) and the Python binding
 * (cudf_amd.groupby) are both thin layers over these entry points.
 *
 * Conventions shared by every entry point
 *   - Columns are described by mgb_column: a type id numbered exactly like
 *     cudf::type_id (cpp/include/cudf/types.hpp:201-233), a device data pointer
 *     to element 0 of the parent buffer, an element offset, a row count and an
 *     optional validity bitmask (32-bit words, LSB first, bit = 1 means valid,
 *     bit index = offset + row), as in cudf::column_view
 *     (cpp/include/cudf/column/column_view.hpp:53-225).
 *   - Aggregation kinds are numbered exactly like cudf::aggregation::Kind
 *     (cpp/include/cudf/aggregation.hpp:85-122).
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream). All
 *     device work is ordered on it.
 *   - Every entry point returns an mgb_status. On failure,
 *     mgb_last_error_message() returns a description that stays valid until
 *     the calling thread's next call into the library.
 */
#ifndef MGB_GROUPBY_H
#define MGB_GROUPBY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The C++ layer maps LOGIC and UNSUPPORTED to cudf::logic_error,
 * HIP to cudf::cuda_error and ALLOC to rmm::bad_alloc
 * (cpp/include/cudf/utilities/error.hpp:63-126). */
typedef enum mgb_status {
  MGB_SUCCESS = 0,
  MGB_ERROR_LOGIC = 1,       /* invalid argument: size mismatch, bad type/agg pair */
  MGB_ERROR_UNSUPPORTED = 2, /* valid in cudf but outside this library (e.g. sort-only aggs) */
  MGB_ERROR_HIP = 3,         /* a HIP runtime call failed */
  MGB_ERROR_ALLOC = 4        /* the allocator returned NULL */
} mgb_status;

/* cudf::type_id subset handled here (same numeric values). */
enum {
  MGB_TYPE_EMPTY = 0,
  MGB_TYPE_INT8 = 1, MGB_TYPE_INT16 = 2, MGB_TYPE_INT32 = 3, MGB_TYPE_INT64 = 4,
  MGB_TYPE_UINT8 = 5, MGB_TYPE_UINT16 = 6, MGB_TYPE_UINT32 = 7, MGB_TYPE_UINT64 = 8,
  MGB_TYPE_FLOAT32 = 9, MGB_TYPE_FLOAT64 = 10, MGB_TYPE_BOOL8 = 11,
  MGB_TYPE_TIMESTAMP_DAYS = 12, MGB_TYPE_TIMESTAMP_SECONDS = 13, MGB_TYPE_TIMESTAMP_MILLISECONDS = 14,
  MGB_TYPE_TIMESTAMP_MICROSECONDS = 15, MGB_TYPE_TIMESTAMP_NANOSECONDS = 16,
  MGB_TYPE_DURATION_DAYS = 17, MGB_TYPE_DURATION_SECONDS = 18, MGB_TYPE_DURATION_MILLISECONDS = 19,
  MGB_TYPE_DURATION_MICROSECONDS = 20, MGB_TYPE_DURATION_NANOSECONDS = 21,
  MGB_TYPE_LIST = 24  /* results only: COLLECT_LIST / COLLECT_SET (offsets + child, see mgb_out_column) */
};

/* cudf::aggregation::Kind subset handled here (same numeric values). */
enum {
  MGB_AGG_SUM = 0, MGB_AGG_PRODUCT = 1, MGB_AGG_MIN = 2, MGB_AGG_MAX = 3,
  MGB_AGG_COUNT_VALID = 4, MGB_AGG_COUNT_ALL = 5, MGB_AGG_SUM_OF_SQUARES = 8,
  MGB_AGG_MEAN = 9, MGB_AGG_M2 = 10, MGB_AGG_VARIANCE = 11, MGB_AGG_STD = 12,
  MGB_AGG_MEDIAN = 13, MGB_AGG_QUANTILE = 14, MGB_AGG_ARGMAX = 15, MGB_AGG_ARGMIN = 16,
  MGB_AGG_NUNIQUE = 17, MGB_AGG_NTH_ELEMENT = 18, MGB_AGG_COLLECT_LIST = 21, MGB_AGG_COLLECT_SET = 22
};

/* cudf::null_equality and cudf::nan_equality (cpp/include/cudf/types.hpp), for COLLECT_SET. */
enum { MGB_NULL_EQUAL = 0, MGB_NULL_UNEQUAL = 1 };
enum { MGB_NAN_ALL_EQUAL = 0, MGB_NAN_UNEQUAL = 1 };

/* cudf::interpolation (cpp/include/cudf/types.hpp). */
enum { MGB_INTERP_LINEAR = 0, MGB_INTERP_LOWER = 1, MGB_INTERP_HIGHER = 2, MGB_INTERP_MIDPOINT = 3,
       MGB_INTERP_NEAREST = 4 };

/* cudf::null_policy (cpp/include/cudf/types.hpp). */
enum { MGB_NULL_EXCLUDE = 0, MGB_NULL_INCLUDE = 1 };

/* Read-only column view (cudf::column_view for fixed-width types). */
typedef struct mgb_column {
  int32_t type_id;
  int32_t reserved;
  int64_t size;              /* rows */
  const void* data;          /* device pointer to element 0 of the parent buffer */
  const uint32_t* null_mask; /* device pointer or NULL (no nulls) */
  int64_t offset;            /* first row, in elements; also the first mask bit */
  int64_t null_count;        /* nulls in [offset, offset+size); -1 = unknown */
} mgb_column;

/* One aggregation: kind plus its parameters (the null_policy of COUNT is
 * expressed by COUNT_VALID vs COUNT_ALL). Zero-initialised fields are the
 * defaults except where noted. */
typedef struct mgb_aggregation {
  int32_t kind;
  int32_t ddof;           /* VARIANCE / STD: ddof; NTH_ELEMENT: n (negative: from the end);
                             COLLECT_SET: MGB_NULL_EQUAL / MGB_NULL_UNEQUAL */
  int32_t null_policy;    /* NUNIQUE, NTH_ELEMENT, COLLECT_LIST, COLLECT_SET: MGB_NULL_EXCLUDE /
                             MGB_NULL_INCLUDE (cudf's defaults are EXCLUDE for NUNIQUE and INCLUDE for
                             NTH_ELEMENT, COLLECT_LIST and COLLECT_SET) */
  int32_t interpolation;  /* QUANTILE: MGB_INTERP_*; COLLECT_SET: MGB_NAN_ALL_EQUAL / MGB_NAN_UNEQUAL */
  double quantile;        /* QUANTILE: one quantile in [0, 1] (MEDIAN is 0.5, LINEAR) */
} mgb_aggregation;

/* cudf::groupby::aggregation_request (cpp/include/cudf/groupby.hpp:59-62). */
typedef struct mgb_request {
  mgb_column values;
  const mgb_aggregation* aggs;
  int32_t num_aggs;
  int32_t reserved;
} mgb_request;

/* Device allocator (the role of the `mr` argument of
 * cudf::groupby::groupby::aggregate, and of cudf's current device resource
 * for temporaries). Result buffers come back owned by the caller, who
 * releases them with `deallocate`, one buffer at a time.
 * `allocate_many` (optional, may be NULL; no counterpart in the reference) is
 * asked for all result buffers of one call at once: n sizes in, n pointers out,
 * 0 on success. A caller whose allocator is costly per call (an interpreter
 * callback) sets it; each buffer is still released by its own `deallocate`.
 * `temporaries` (optional, may be NULL: this allocator) serves every device
 * temporary of a call (hash tables, partitioned rows, histograms, staged
 * results, sort space): each is allocated during the call and returned with
 * `deallocate` on the call's stream before the call returns. The return is
 * stream-ordered, as rmm's resources are: work already queued on that stream
 * may still use the buffer, so the allocator hands it to other streams only
 * after that work (an allocator that frees at once must wait for the stream
 * first). Nothing else of a call touches device memory the caller did not
 * provide. */
typedef struct mgb_allocator {
  void* (*allocate)(void* ctx, size_t bytes, void* stream);
  void (*deallocate)(void* ctx, void* ptr, size_t bytes, void* stream);
  void* ctx;
  int32_t (*allocate_many)(void* ctx, int32_t n, const size_t* bytes, void** out, void* stream);
  const struct mgb_allocator* temporaries;
} mgb_allocator;

/* An owned result column written by the library. A LIST column
 * (MGB_TYPE_LIST, the results of COLLECT_LIST and COLLECT_SET; cudf's lists
 * column, cpp/include/cudf/lists/lists_column_view.hpp) keeps its size + 1
 * int32 offsets in `data` and never has nulls itself; its elements are the
 * child column described by the child_* fields (zero for other columns). */
typedef struct mgb_out_column {
  int32_t type_id;
  int32_t reserved;
  int64_t size;
  void* data;          /* allocated with the caller's allocator; NULL when size == 0 */
  uint32_t* null_mask; /* allocated with the caller's allocator, or NULL */
  int64_t null_count;
  size_t data_bytes;   /* bytes allocated for data (for deallocate) */
  size_t mask_bytes;   /* bytes allocated for null_mask */
  int32_t child_type_id;
  int32_t child_reserved;
  int64_t child_size;
  void* child_data;          /* allocated with the caller's allocator; NULL when child_size == 0 */
  uint32_t* child_null_mask; /* allocated with the caller's allocator, or NULL */
  int64_t child_null_count;
  size_t child_data_bytes;
  size_t child_mask_bytes;
} mgb_out_column;

/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int32_t mgb_version(void);

/* Message of the calling thread's most recent failure ("" if none). */
const char* mgb_last_error_message(void);

/* Result type id of (value type, aggregation kind), or -1 when the pair is not
 * a valid aggregation (cudf::detail::target_type / is_valid_aggregation,
 * cpp/include/cudf/detail/aggregation/aggregation.hpp:1168-1424,1610-1623). */
int32_t mgb_target_type(int32_t value_type_id, int32_t agg_kind);

/*
 * Hash groupby-aggregate. Replaces cudf::groupby::groupby::aggregate
 * (cpp/include/cudf/groupby.hpp:185-197, cpp/src/groupby/groupby.cu:195-219,
 * hash path cpp/src/groupby/hash/groupby.cu:548-665).
 *
 *   keys, num_keys        the key table (all columns the same size)
 *   null_policy           MGB_NULL_EXCLUDE drops rows with any null key;
 *                         MGB_NULL_INCLUDE groups null keys together
 *   requests              value columns and their aggregations
 *   out_keys              num_keys entries: the unique key rows
 *   out_results           one entry per aggregation, requests in order and
 *                         each request's aggregations in order
 *   out_num_groups        number of groups (rows of every output)
 *
 * Output row order is arbitrary (as in cudf). The call waits on `stream` for
 * the group count; the outputs' last writes may still be queued on `stream`
 * when it returns (stream order, as in cudf): a reader on another stream
 * waits for `stream` first. Errors: MGB_ERROR_LOGIC for a
 * size mismatch or an invalid type/aggregation pair (groupby.cu:208-212,
 * 173-190); MGB_ERROR_UNSUPPORTED for more rows than cudf::size_type holds
 * and for the aggregations of nested values or nested results other than
 * COLLECT_LIST / COLLECT_SET (MERGE_*, TDIGEST, HISTOGRAM, ...);
 * MGB_ERROR_ALLOC when the allocator returns NULL. VARIANCE, STD, ARGMIN and
 * ARGMAX (hash aggregations in cpp/src/groupby/hash/groupby.cu:73-84) and M2
 * (sort-only in cudf, cpp/src/groupby/sort/group_m2.cu) take a second pass
 * over the rows; MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT, COLLECT_LIST and
 * COLLECT_SET (sort-only in cudf, cpp/src/groupby/sort/aggregate.cpp) sort
 * each value column's rows by group. COLLECT_LIST gives each group's values
 * in row order (null values as null elements, or dropped under
 * MGB_NULL_EXCLUDE); COLLECT_SET its distinct values in ascending order
 * (NaN after every number, -0.0 == +0.0 with -0.0 kept when present, NaNs
 * one element or each its own under MGB_NAN_UNEQUAL), then its nulls (one
 * under MGB_NULL_EQUAL, each under MGB_NULL_UNEQUAL, none under
 * MGB_NULL_EXCLUDE). Both are LIST columns of the value type (see
 * mgb_out_column); calls with them wait on `stream` once more, for their
 * element counts.
 */
mgb_status mgb_groupby_aggregate(const mgb_column* keys, int32_t num_keys, int32_t null_policy,
                                 const mgb_request* requests, int32_t num_requests,
                                 mgb_out_column* out_keys, mgb_out_column* out_results,
                                 int64_t* out_num_groups, void* stream, const mgb_allocator* alloc);

/*
 * mgb_groupby_aggregate for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. Same inputs, semantics and
 * errors as mgb_groupby_aggregate, but the call allocates nothing: its
 * temporaries come from a fixed 256 MiB scratch array the library holds on
 * each device it runs on (for the life of the process), and the results go
 * into buffers the caller provides:
 *   out_keys[c] / out_results[j]   data: a device buffer of data_bytes bytes
 *                                  holding max_groups values of the column's
 *                                  type (mgb_target_type for results);
 *                                  null_mask: a device buffer of mask_bytes
 *                                  bytes holding max_groups bits (rounded up
 *                                  to 64 bytes), needed where the column can
 *                                  hold nulls and set to NULL on return where
 *                                  it has none. type_id, size and null_count
 *                                  are written by the call.
 * Inputs above 8,388,608 rows are refused (MGB_ERROR_UNSUPPORTED) before any
 * device work; a call whose temporaries would pass the scratch array, or
 * whose group count passes max_groups, fails with MGB_ERROR_ALLOC. Calls of
 * one process run one at a time (they share the scratch array) and return
 * with their work on `stream` complete.
 */
mgb_status mgb_groupby_aggregate_noalloc(const mgb_column* keys, int32_t num_keys, int32_t null_policy,
                                         const mgb_request* requests, int32_t num_requests,
                                         mgb_out_column* out_keys, mgb_out_column* out_results,
                                         int64_t max_groups, int64_t* out_num_groups, void* stream);

/*
 * Row hash of a table with cudf's MurmurHash3_x86_32 row hasher: per column
 * murmur3 of the element (floats with -0.0 -> +0.0 and NaN -> canonical NaN,
 * null -> 0xFFFFFFFF), folded with hash_combine from `seed`. Replaces
 * cudf::hashing::murmurhash3_x86_32(table_view, seed) for fixed-width columns
 * (cpp/include/cudf/hashing.hpp:90-94, cpp/src/hash/murmurhash3_x86_32.cu).
 * `out` is a device array of num_rows uint32.
 */
mgb_status mgb_murmurhash3_x86_32(const mgb_column* columns, int32_t num_columns, uint32_t seed,
                                  uint32_t* out, void* stream);

/* Estimated number of distinct key rows (groups) of a key table: exact up
 * to 131072 rows, otherwise from a sample, or HyperLogLog over every row when
 * the sample does not settle it (the plain path's own estimate). Its
 * temporaries (about 40 MB at most) come from `alloc` (its `temporaries` if
 * set). Synchronises `stream`. No counterpart in the reference. */
mgb_status mgb_estimate_groups(const mgb_column* keys, int32_t num_keys, int32_t null_policy, void* stream,
                               const mgb_allocator* alloc, int64_t* estimate);

/*
 * Spark-compatible row hash: columns hashed in order with each result seeding
 * the next, nulls leaving the running hash unchanged, small integers widened
 * to 4 bytes. Replaces cudf::hashing::spark_murmurhash3_x86_32(table_view,
 * seed) for fixed-width columns (cpp/include/cudf/hashing.hpp:130-134,
 * cpp/src/hash/spark_murmurhash3_x86_32.cu). `out` is num_rows int32.
 */
mgb_status mgb_spark_murmurhash3_x86_32(const mgb_column* columns, int32_t num_columns, uint32_t seed,
                                        int32_t* out, void* stream);

/*
 * Measurement hooks (no counterpart in cudf). When enabled, every launch of
 * a named kernel ("aggregate", "fill", "finalize") is bracketed with HIP
 * events on its stream. mgb_profiling_read synchronises those events and
 * returns the number of launches and their summed duration since the last
 * read, then forgets them.
 */
mgb_status mgb_profiling_enable(int32_t enable);
mgb_status mgb_profiling_read(const char* kernel, int64_t* launches, double* total_ms);

/* Releases the pinned host staging the library keeps for `stream` (readback
 * slots and parameter uploads, about 72 KB, taken at the stream's first call).
 * Call after the stream's last groupby call, e.g. before destroying it. */
mgb_status mgb_release_workspace(void* stream);

/* ------------------------------------------------------------------------
 * Device memory pool (no counterpart in cudf; the role of rmm's
 * pool_memory_resource). Chunks are taken from the device as needed, never
 * more than `limit_bytes` in all, and kept until the pool is trimmed or
 * destroyed; blocks are carved first-fit. A block returned on a stream is
 * reused at once by that stream and by other streams only after the pool has
 * waited for it (which it does only when it would otherwise pass its limit).
 * mgb_pool_allocator fills an mgb_allocator over the pool, e.g. as the
 * `temporaries` of a groupby call's allocator.
 * ------------------------------------------------------------------------ */
typedef struct mgb_pool mgb_pool;

typedef struct mgb_pool_stats {
  int64_t limit_bytes;
  int64_t held_bytes;         /* chunks taken from the device */
  int64_t in_use_bytes;       /* blocks handed out and not returned */
  int64_t peak_in_use_bytes;
  int64_t chunks;
  int64_t stream_waits;       /* streams waited for to reuse their blocks elsewhere */
} mgb_pool_stats;

mgb_status mgb_pool_create(int64_t limit_bytes, mgb_pool** out);
mgb_status mgb_pool_allocator(mgb_pool* pool, mgb_allocator* out);
mgb_status mgb_pool_get_stats(mgb_pool* pool, mgb_pool_stats* out);
/* Waits for the streams that returned blocks and frees the chunks no block uses. */
mgb_status mgb_pool_trim(mgb_pool* pool);
/* Waits for the streams that returned blocks, then frees every chunk. Every
 * block must have been returned. */
void mgb_pool_destroy(mgb_pool* pool);


/* ------------------------------------------------------------------------
 * Batched execution (no counterpart in the reference: cudf runs one
 * groupby per call). Many independent groupby-aggregate queries that share
 * one schema -- the same key types, value types, null-ness and aggregations
 * -- are planned together and run as one HIP graph: per query the planner
 * estimates the number of groups from a sample and picks the LDS-direct,
 * global-atomic or hash-partitioned execution, the LDS table size and the
 * work-item size by a cost model; the work items of all queries are then
 * taken from one queue by a persistent kernel.
 *
 * Lifecycle: mgb_batch_create (plans; synchronises `stream` once), then any
 * number of mgb_batch_execute (enqueues the graph; does not wait) each
 * followed by mgb_batch_finish (waits, reads the group counts, re-runs on
 * the plain path any query whose estimate proved too low), then
 * mgb_batch_result per query, then mgb_batch_destroy. The input columns
 * must stay valid and unchanged for the batch's lifetime. Results are views
 * into buffers the batch owns, valid until the next execute or destroy.
 * One batch must not be executed on two streams at once.
 * ------------------------------------------------------------------------ */
typedef struct mgb_batch mgb_batch;

typedef struct mgb_batch_options {
  int32_t force_mode;        /* -1: chosen per query; 0: LDS direct; 1: global atomics; 2: partitioned */
  int32_t use_graph;         /* 1: capture the execution in a HIP graph; 0: plain launches */
  int64_t direct_item_rows;  /* rows per direct / global work item; 0: chosen per query */
  double estimate_scale;     /* multiplies every distinct-count estimate (1.0; below 1 only to test re-runs) */
  const int64_t* estimates;  /* distinct key rows per query (e.g. cached from mgb_estimate_groups for
                                resident tables), or NULL: sampled on the device (one synchronisation) */
  const mgb_allocator* alloc; /* the batch's device memory (one allocation, freed by mgb_batch_destroy on
                                 the create stream), or NULL: stream-ordered from the device's pool */
} mgb_batch_options;

typedef struct mgb_batch_stats {
  int64_t num_queries;
  int64_t direct_queries;
  int64_t global_queries;
  int64_t partitioned_queries;
  int64_t num_items;             /* work items of the persistent kernel */
  int64_t output_rows_reserved;  /* rows of the concatenated output columns */
  int64_t device_bytes;          /* device memory held by the batch */
  int64_t reruns;                /* queries re-run on the plain path by the last finish */
} mgb_batch_stats;

/* keys: num_queries * num_keys columns, query-major; requests:
 * num_queries * num_requests, query-major (request r of every query carries
 * the same aggregations). options may be NULL. MGB_ERROR_UNSUPPORTED for
 * VARIANCE, STD, M2, ARGMIN, ARGMAX and for null keys grouped under
 * MGB_NULL_INCLUDE: those queries run on mgb_groupby_aggregate. */
mgb_status mgb_batch_create(const mgb_column* keys, int32_t num_keys, int32_t null_policy,
                            const mgb_request* requests, int32_t num_requests, int32_t num_queries,
                            const mgb_batch_options* options, void* stream, mgb_batch** out);
mgb_status mgb_batch_execute(mgb_batch* batch, void* stream);
mgb_status mgb_batch_finish(mgb_batch* batch, void* stream);
/* Views of query `query`'s result: out_keys[num_keys], out_results[one per
 * aggregation, in request order]. A view's rows start at its `offset`
 * (element offset of data, bit offset of null_mask); null_count is -1 when
 * not counted. */
mgb_status mgb_batch_result(const mgb_batch* batch, int32_t query, mgb_column* out_keys, mgb_column* out_results,
                            int64_t* num_groups);
/* Query `query`'s result copied into caller-owned columns (allocated with
 * `alloc`, validity re-packed from bit 0, null counts exact), as
 * mgb_groupby_aggregate returns them. Synchronises `stream`. */
mgb_status mgb_batch_copy_result(const mgb_batch* batch, int32_t query, mgb_out_column* out_keys,
                                 mgb_out_column* out_results, int64_t* num_groups, void* stream,
                                 const mgb_allocator* alloc);
mgb_status mgb_batch_get_stats(const mgb_batch* batch, mgb_batch_stats* out);
/* The execution chosen for a query (0 direct, 1 global, 2 partitioned) and
 * its distinct-count estimate. */
mgb_status mgb_batch_query_mode(const mgb_batch* batch, int32_t query, int32_t* mode, int64_t* estimate);
void mgb_batch_destroy(mgb_batch* batch);


/* ------------------------------------------------------------------------
 * Serving (no counterpart in the reference). A server keeps tables resident
 * on the device and answers groupby queries on them submitted from any host
 * thread. A scheduler thread batches the queries that are waiting
 * (continuous batching: whatever arrived while earlier batches ran), admits
 * each batch against the HBM budget (resident tables + results not yet taken
 * + the batch's own memory; a batch that does not fit is halved, a query
 * that does not fit waits for memory to be released), and runs up to
 * max_inflight batches at once, one stream each, so their device work
 * overlaps. Batches and results take their memory from the server's own
 * device chunks, allocated at creation (reserve_bytes) or when first needed
 * and kept until mgb_server_destroy: new device memory is slow to map and
 * clear, and would stall the work in flight. With batching = 0 every query runs alone on the plain path, in
 * arrival order (the per-query baseline). Results do not depend on which
 * batch a query ran in, beyond floating-point summation order.
 * ------------------------------------------------------------------------ */
typedef struct mgb_server mgb_server;

typedef struct mgb_server_options {
  int64_t hbm_budget_bytes;   /* 0: 200 GiB */
  int32_t max_inflight;       /* batches executing at once (0: 2) */
  int32_t max_batch_queries;  /* queries per batch at most (0: 256) */
  int32_t batching;           /* 1: continuous batching; 0: per-query plain path */
  int32_t reserved;
  int64_t reserve_bytes;      /* device memory for batches and results taken at creation (0: none);
                                 the server grows it on demand up to the budget and keeps it */
} mgb_server_options;

/* One request of a query: a column of the table and its aggregations. */
typedef struct mgb_server_request {
  int32_t column;
  int32_t num_aggs;
  const mgb_aggregation* aggs;
} mgb_server_request;

typedef struct mgb_server_stats {
  int64_t submitted;
  int64_t completed;
  int64_t failed;
  int64_t batches;
  int64_t deferred;           /* batches halved or postponed by the HBM budget */
  int64_t reruns;             /* queries re-run on the plain path after an underestimate */
  int64_t peak_held_bytes;
  int64_t resident_bytes;
  int64_t held_bytes;
  int64_t pending;
  double mean_batch_queries;
  double host_plan_ms;        /* scheduler time spent planning batches */
  double host_finish_ms;      /* scheduler time spent finishing batches */
  int64_t device_bytes;       /* device memory the server holds for batches and results */
} mgb_server_stats;

mgb_status mgb_server_create(const mgb_server_options* options, mgb_server** out);
/* Registers a device-resident table; its columns stay owned by the caller and
 * must outlive the server. Returns MGB_ERROR_ALLOC past the budget. */
mgb_status mgb_server_add_table(mgb_server* server, const mgb_column* columns, int32_t num_columns,
                                int32_t* table_id);
/* Queues a query: group table `table_id` by columns key_columns[num_keys]. */
mgb_status mgb_server_submit(mgb_server* server, int32_t table_id, const int32_t* key_columns, int32_t num_keys,
                             int32_t null_policy, const mgb_server_request* requests, int32_t num_requests,
                             int64_t* ticket);
/* Blocks until the ticket's query has finished (timeout_ms < 0: no limit);
 * reports its group count and its latency from submission to completion. */
mgb_status mgb_server_wait(mgb_server* server, int64_t ticket, double timeout_ms, int64_t* num_groups,
                           double* latency_ms);
/* Copies a finished query's result into caller-owned columns (allocated with
 * `alloc`, as mgb_groupby_aggregate returns them) and releases the ticket. */
mgb_status mgb_server_take_result(mgb_server* server, int64_t ticket, mgb_out_column* out_keys,
                                  mgb_out_column* out_results, int64_t* num_groups, void* stream,
                                  const mgb_allocator* alloc);
/* Releases a finished ticket without copying its result. */
mgb_status mgb_server_release(mgb_server* server, int64_t ticket);
mgb_status mgb_server_get_stats(mgb_server* server, mgb_server_stats* out);
/* Stops the scheduler after the batches in flight; queries still queued are
 * dropped. Call after every wait has returned. */
void mgb_server_destroy(mgb_server* server);

#ifdef __cplusplus
}
#endif

#endif /* MGB_GROUPBY_H */
