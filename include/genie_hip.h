/*
 * genie_hip.h -- C ABI of the MI355X match-count top-k engine.
 *
 * Everything below GENIE's matching::knn (src/genie/matching/knn.cu:106-183)
 * is replaced by these entry points: the host library (genie_amd/csrc/host,
 * which keeps GENIE's table-build / query / top-k C++ API) calls only this
 * header.  Plain pointers and sizes, int status codes, no C++ or torch types.
 *
 * Status codes: 0 ok; GENIE_E_RUNTIME (HIP error), GENIE_E_OOM (device
 * allocation failed), GENIE_E_INVALID (bad argument), GENIE_E_NOT_BUILT
 * (index handle empty / wrong kind).  genie_hip_last_error() returns a
 * thread-local message for the last failure on the calling thread.
 * The host library maps them onto the reference's exception classes
 * (src/genie/exception/exception.h:10-46).
 */
#ifndef GENIE_HIP_H_
#define GENIE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GENIE_OK 0
#define GENIE_E_RUNTIME 1
#define GENIE_E_OOM 2
#define GENIE_E_INVALID 3
#define GENIE_E_NOT_BUILT 4

/* Largest k per query any entry point accepts. */
#define GENIE_MAX_K 4096

/* Work item, byte-for-byte the reference's Query::dim (src/genie/query/
 * query.h:59-68): 20 bytes, [start_pos, end_pos) into the index's flat
 * posting array (uncompressed) or compressed word array (compressed). */
typedef struct genie_dim_t {
    int32_t query;
    int32_t order;
    int32_t start_pos;
    int32_t end_pos;
    float weight;
} genie_dim_t;

/* Description of a built index.  Mirrors what inv_table / inv_compr_table
 * hold after build() (src/genie/table/inv_table.h:110-174,
 * inv_compr_table.cu:31-119), re-expressed with a dense key numbering:
 * key j of dim d is key_base[d] + (value - lo[d]); key_start[j] is the
 * position in the uncompressed posting array where that key's ids begin
 * (n_keys + 1 entries, last = n_postings).  Segments are the _inv_pos splits
 * (seg_upos, n_segments + 1 entries); with codec != 0 the words array holds
 * each segment compressed independently and seg_cpos (n_segments + 1) gives
 * their word offsets -- the positions compressed work items refer to. */
typedef struct genie_index_desc {
    uint32_t n_objects;          /* inv_table::i_size() */
    uint32_t n_dims;             /* inv_table::m_size() */
    uint32_t n_keys;
    uint64_t n_postings;
    const uint32_t* postings;    /* [n_postings] ids, key-major, ascending per key (may be NULL when codec != 0) */
    const uint64_t* key_start;   /* [n_keys + 1] */
    const uint32_t* key_dim;     /* [n_keys] dim of each key */
    uint32_t n_segments;
    const uint32_t* seg_upos;    /* [n_segments + 1] uncompressed positions (_inv_pos) */
    int32_t codec;               /* COMPRESSION_TYPE (DeviceCodecFactory.h:38-53); 0 = uncompressed */
    uint64_t n_words;
    const uint32_t* words;       /* [n_words] compressed segments (codec != 0) */
    const uint32_t* seg_cpos;    /* [n_segments + 1] word offsets (codec != 0) */
    int32_t single_value;        /* 1: every object holds exactly one value per dim */
} genie_index_desc;

typedef struct genie_index genie_index;

/* Select the device (Init, src/genie/utility/init.cu:7-34; an out-of-range id
 * falls back to device 0 as the reference does). */
int genie_hip_init(int device);

/* Upload an index and keep it resident in HBM (inv_table::cpy_data_to_gpu,
 * inv_table.cu:30-45; inv_compr_table::cpy_data_to_gpu, inv_compr_table.cu:
 * 211-223), then build the per-key tile directory on the device. */
int genie_hip_index_create(const genie_index_desc* desc, genie_index** out);
int genie_hip_index_destroy(genie_index* ix);
/* Bytes of HBM the handle holds. */
uint64_t genie_hip_index_bytes(const genie_index* ix);

/* Upper bound on any query's count, from the items (weights must be
 * non-negative integers).  out_bounds[q] for q < n_queries. */
int genie_hip_count_bounds(const genie_index* ix, const genie_dim_t* items, uint64_t n_items,
                           uint32_t n_queries, uint32_t* out_bounds);

/* Per-batch search: replaces knn() -> match()/match_integrated() ->
 * heap_count_topk() -> extract_index_and_count() (knn.cu:106-183,
 * match.cu:224-464, heap_count.cu:161-282, knn.cu:48-55).  Runs as an
 * automatic plan (mode 2): the dense form when the batch qualifies, the
 * plain form otherwise; the results are the same either way.  items in HOST
 * memory (any order; item.query in [0, n_queries)); k_per_query[q] in
 * [1, GENIE_MAX_K]; count_bound 0 = derive from the items.  Output is
 * query-major with out_offsets[q] = sum of k over earlier queries, k_q slots
 * per query, ordered count descending then id ascending, padded with
 * (id 0, count 0) (heap_count.cu:254-255).  Synchronous. */
int genie_hip_match_topk(genie_index* ix, const genie_dim_t* items, uint64_t n_items,
                         uint32_t n_queries, const uint32_t* k_per_query, uint32_t count_bound,
                         int32_t* out_ids, int32_t* out_counts);
/* The same search ordered on the caller's stream (a hipStream_t, NULL = the
 * default stream): its copies and kernels are issued on `stream` after the
 * caller's earlier work there, and the call returns once the results are in
 * out_ids / out_counts (host or device memory), waiting for `stream` only.
 * genie_hip_match_topk is this call on the default stream.  For a call that
 * returns before the results are ready, use a plan: genie_hip_plan_run and
 * genie_hip_plan_download_async on the caller's stream. */
int genie_hip_match_topk_stream(genie_index* ix, const genie_dim_t* items, uint64_t n_items,
                                uint32_t n_queries, const uint32_t* k_per_query, uint32_t count_bound,
                                int32_t* out_ids, int32_t* out_counts, void* stream);

/* ---- device-resident form (inputs already in HBM) ---------------------- */
/* A plan holds device scratch for batches up to the given shape, so the hot
 * call allocates nothing.  mode: 0 plain (one query per workgroup),
 * 1 batched (the dense form when the batch qualifies -- see
 * genie_hip_plan_set_dense -- else queries packed into count lanes over a
 * persistent work queue), 2 automatic (the dense form when the batch
 * qualifies, else plain).  All modes give identical results. */
typedef struct genie_plan genie_plan;
int genie_hip_plan_create(genie_index* ix, uint32_t max_queries, uint64_t max_items,
                          uint32_t max_k, uint32_t count_bound, int mode, genie_plan** out);
int genie_hip_plan_destroy(genie_plan* p);
/* Device memory a plan of this shape holds for one batch of n_queries queries
 * and n_items items with k <= max_k (its candidate space as first sized, the
 * dense form's tables, the full-layout decode scratch of a compressed
 * index's plain form): the admission estimate of knn_MT (the reference's
 * estimate is knn.cu:218-227). */
int genie_hip_plan_estimate(const genie_index* ix, uint32_t n_queries, uint64_t n_items, uint32_t max_k, int mode,
                            uint64_t* bytes);
/* Take a batch (host arrays; they may be reused on return): its device
 * copies are staged in the plan's pinned host buffer and issued by the next
 * genie_hip_plan_run on that run's stream, ahead of its kernels. */
int genie_hip_plan_upload(genie_plan* p, const genie_dim_t* items, uint64_t n_items,
                          uint32_t n_queries, const uint32_t* k_per_query);
/* Run the search on the uploaded batch on `stream` (a hipStream_t, NULL =
 * default stream); asynchronous. */
int genie_hip_plan_run(genie_plan* p, void* stream);
/* Copy the results of the last run to the caller's buffers (host or device
 * memory), on `stream`, and wait for them (synchronises stream). */
int genie_hip_plan_download(genie_plan* p, void* stream, int32_t* out_ids, int32_t* out_counts);
/* Per-stage timing of the last run when timing is on (HIP events recorded
 * on the run's stream): "compile" (item -> chunk kernels; query masks for
 * the dense form), "sample" (the dense form's sample pass and thresholds,
 * 0 otherwise), "scan" (the count + per-tile select kernel; the dense main
 * pass), "merge"; ms is the stage's duration.  "decode": the part of
 * "scan" that decoded a compressed batch's segments into a scratch (0 when
 * the run decoded none). */
int genie_hip_plan_set_timing(genie_plan* p, int on);
int genie_hip_plan_kernel_ms(genie_plan* p, const char* name, double* ms);
/* Replay the run as a captured hipGraph (captured at the next run, again
 * after any upload or configuration change).  Timing runs bypass the graph. */
int genie_hip_plan_set_graph(genie_plan* p, int on);
/* Batched-form configuration: threads per workgroup (256/512/1024), objects
 * per tile (a multiple of 8192 whose 8-byte counters fit LDS) and the
 * longest posting span one wave takes at a time. */
int genie_hip_plan_set_config(genie_plan* p, uint32_t block, uint32_t tile_objs, uint32_t seglen);
int genie_hip_plan_get_config(genie_plan* p, uint32_t* block, uint32_t* tile_objs, uint32_t* seglen,
                              uint32_t* lane_bits);
/* Autotuner of a batched plan on its uploaded batch: the lane form over block
 * x tile x segment length (12 candidates), then the dense or set form when
 * the batch allows it and it is switched on, then the plain form, so that a
 * batched plan never runs slower than the plain one; keeps the fastest (form
 * and configuration).  out_ms (optional, >= 14 entries) receives every
 * candidate's time per run in that order (-1 if it could not run); *n_out the
 * number of candidates. */
int genie_hip_plan_autotune(genie_plan* p, int reps, double* out_ms, uint32_t* n_out);
/* Each query's adaptive threshold after the last run (synchronises stream;
 * out[n_queries]), the role of updateThreshold's d_threshold
 * (match_device_utils.h:449-473).  Tile forms: the largest k-th count of any
 * tile that holds at least k objects counting at least 1 (0 if none); with
 * one tile over the whole index that is the k-th largest count c_(k), the
 * reference's closed form thr - 1 = c_(k).  Dense and set forms: the
 * threshold their main pass emitted from.  Measurement and tests. */
int genie_hip_plan_thresholds(genie_plan* p, void* stream, uint32_t* out);
/* Batched plans: the form of batches the dense and set forms do not take:
 * 0 the plain form, 1 the lane form (the default until an autotune). */
int genie_hip_plan_set_form(genie_plan* p, int form);
/* The form the uploaded batch runs: 0 plain, 1 lane, 2 dense, 3 set. */
int genie_hip_plan_form(genie_plan* p, int* form);
/* Dense form of the batched search (modes 1 and 2; on by default).  When the
 * index is a single-value table of at most 128 dims with at most 256 keys per
 * dim, and every item of the uploaded batch covers whole keys with an integer
 * weight from 1 to 15 (never one key twice for a query), the batch is counted
 * from per-object key codes built at index creation (a compressed table is
 * decoded once on the device for this), 32 queries per LDS lookup with
 * bit-sliced counters (8 queries of four weight bits each when a weight is
 * above 1), instead of from the posting lists; results are identical.
 * The same switch covers the set form of set-valued (bijectMap) tables: when
 * each query's weight per key is an integer up to 63 (items summed over the
 * positions they cover, constant over each key), the longest forward list
 * times the query's largest weight is below 2047, and each group of 32
 * queries names at most 2048 keys, objects are counted over their forward
 * lists (built at index creation) against an LDS hash table of the group's
 * keys, with 11 bit-sliced count planes.
 * No counterpart in the reference (its matching always walks the lists,
 * match.cu:161-220).  on = 0 forces the lane form. */
int genie_hip_plan_set_dense(genie_plan* p, int on);
/* Threads per workgroup of the dense and set forms' kernels: 0 (default)
 * runs 1,024 when the runtime's occupancy query places such a workgroup and
 * 512 otherwise (some runtimes place no 1,024-thread workgroup of more than
 * 64 VGPRs); 512 or 1024 forces one (1024 then fails where it is not
 * placed).  Results are identical.  No counterpart in the reference. */
int genie_hip_plan_set_dense_threads(genie_plan* p, int threads);
/* Whether the last uploaded batch runs the dense form (or the set form). */
int genie_hip_plan_dense_active(genie_plan* p, int* active);
/* Sparse select of the plain form: counting marks the counter words a
 * (query, tile) touches and the per-tile top-k visits only those.  mode -1
 * (default) chooses it when the batch averages fewer postings per (query,
 * tile) than a quarter of the tile's counter words; 0 never, 1 always.  A
 * tile that touched more than an eighth of its counter words takes the
 * dense select either way.  Results are identical.  The reference has no
 * counterpart: its per-query c-PQ hash tables hold only the ids a query
 * touches (match.cu:332-341, grown on overflow at 397-404). */
int genie_hip_plan_set_sparse(genie_plan* p, int mode);
/* Whether the plain form of the last uploaded batch selects sparsely. */
int genie_hip_plan_sparse_active(genie_plan* p, int* active);
/* Hash form of the plain scan (modes 0 and 2).  A selective query (mode -1,
 * the default: at most one posting per 16 objects; mode 1: every query whose
 * counts fit) is counted per tile of about 8,192 of its postings in an LDS
 * open-addressing table instead of one counter per object, and the tile's
 * top k taken by an exact radix select.  A query whose table fills up is run
 * on the direct-mapped form at download.  mode 0: never.  Results are
 * identical in every mode.  It replaces the reference's per-query c-PQ hash
 * table (match_device_utils.h:161-419), which evicts entries and can lose
 * the k-th count; these tables keep every touched id. */
int genie_hip_plan_set_hash(genie_plan* p, int mode);
/* Decode of a compressed index in the plain form (modes 0 and 2).  mode -1
 * (default): the segments the uploaded batch names are decoded once each per
 * run into a plan-owned scratch of the uncompressed layout (4 B per posting
 * of the index, allocated at first use), and the batch is counted from it by
 * the uncompressed kernels; where that scratch cannot be allocated, or with
 * mode 0, the kernels decode segments themselves (the direct-mapped kernel
 * once per (query, tile) it overlaps; the hash form once per query item into
 * a packed scratch).  Results are identical.  The reference decodes each
 * (query, segment) work item in shared memory inside its match kernel
 * (match_integrated.cu:169-325). */
int genie_hip_plan_set_decode(genie_plan* p, int mode);
/* Whether the last uploaded batch decodes its segments once into the
 * full-layout scratch. */
int genie_hip_plan_decode_active(genie_plan* p, int* active);
/* Number of queries of the last uploaded batch on the hash form. */
int genie_hip_plan_hash_queries(genie_plan* p, uint32_t* n);
/* Host time of the last batched work-queue build (upload), ms. */
int genie_hip_plan_host_ms(genie_plan* p, double* ms);
/* Algorithmic bytes of the last uploaded batch: logical = 4 B x sum of
 * item span lengths (uncompressed) or words (compressed); compulsory =
 * the same over distinct keys/segments touched, counted once. */
int genie_hip_plan_bytes(genie_plan* p, uint64_t* logical, uint64_t* compulsory);

/* Asynchronous result copy into caller-owned pinned host memory
 * (genie_hip_host_alloc): ids, counts and the rerun flag.  The flag reads
 * nonzero after the stream completes when the candidate space overflowed or a
 * hash-form query's table filled up; the rows are then incomplete: call
 * genie_hip_plan_download, which grows the candidate space or moves those
 * queries to the direct-mapped kernel, reruns and copies the final rows. */
int genie_hip_plan_download_async(genie_plan* p, void* stream, int32_t* out_ids, int32_t* out_counts,
                                  uint32_t* overflow);
int genie_hip_host_alloc(void** ptr, uint64_t bytes);
int genie_hip_host_free(void* ptr);

/* ---- serving form ------------------------------------------------------- */
/* A request loop over a resident index: requests (the work items of ONE
 * query plus its k) are admitted while the estimated count-table memory of
 * the in-flight ones stays within the budget, packed into batches as they
 * arrive and run on several executors (plan + stream each), so batches that
 * arrive while others are on the device overlap with them.  Results are per
 * request, identical to running the request alone. */
typedef struct genie_server_config {
    uint32_t executors;       /* concurrent batches on the device (streams), >= 1 */
    uint32_t max_batch;       /* queries per batch */
    uint64_t max_batch_items; /* work items per batch */
    uint32_t max_k;           /* largest k accepted */
    uint64_t memory_budget;   /* bytes of device scratch for in-flight requests */
    int mode;                 /* plan mode of the executors: 0 plain, 1 batched, 2 automatic */
} genie_server_config;

typedef struct genie_server genie_server;
int genie_hip_server_create(genie_index* ix, const genie_server_config* cfg, genie_server** out);
int genie_hip_server_destroy(genie_server* s);
/* Thread-safe.  items: the query's work items (their .query is ignored). */
int genie_hip_server_submit(genie_server* s, const genie_dim_t* items, uint32_t n_items, uint32_t k,
                            uint64_t* ticket);
/* Block until the request is done; copies its k results and its latency
 * (submit to completion, ms).  Each ticket can be waited once. */
int genie_hip_server_wait(genie_server* s, uint64_t ticket, int32_t* out_ids, int32_t* out_counts,
                          double* latency_ms);
/* Open-loop trace replay: request r (items [item_ptr[r], item_ptr[r+1]),
 * k_per_request[r]) is submitted at arrival_s[r] seconds after the start by
 * one of 4 generator threads (r mod 4); returns per-request latency (ms, from the scheduled
 * arrival to completion, so a generator running late counts in it), all
 * results (out_off[r] = sum of earlier k), and the wall time of the whole
 * replay. */
int genie_hip_server_replay(genie_server* s, const genie_dim_t* items, const uint64_t* item_ptr,
                            const uint32_t* k_per_request, const double* arrival_s, uint32_t n_requests,
                            double* out_latency_ms, int32_t* out_ids, int32_t* out_counts, double* wall_s);
/* Batches run and requests completed so far. */
int genie_hip_server_stats(genie_server* s, uint64_t* batches, uint64_t* requests, uint64_t* max_inflight);
/* Seconds the executors spent so far, summed over executors, per phase of a
 * batch: out[0] gathering the batch and uploading it (host work and copies),
 * out[1] launching its kernels and result copies, out[2] waiting for them,
 * out[3] handing the results back; out[4] the number of batches run again
 * after their candidates overflowed the space first given them (out holds 5
 * values).  Measurement only (no counterpart in the reference, whose
 * dist_genie scheduler has no timing). */
int genie_hip_server_phase_seconds(genie_server* s, double* out);
/* The longest single batch per phase (out[0..3], as above, over executors)
 * and the longest time a request waited between its submission and its batch
 * going to an executor (out[4]); out holds 5 values.  reset != 0 clears them
 * after reading, so that each call covers the span since the last one.
 * Measurement only (no counterpart in the reference). */
int genie_hip_server_phase_max_seconds(genie_server* s, double* out, int reset);

/* ---- no-allocation entry point ------------------------------------------ */
/* genie_hip_match_topk without device allocation, for callers that cannot
 * allocate (code that must not call hipMalloc: tight serving loops,
 * harnesses that forbid it).  No counterpart in the reference.  Same
 * arguments and results as genie_hip_match_topk (host arrays in and out,
 * returns when the results are copied), run on `stream` (a hipStream_t,
 * NULL = default stream) as the plain form's direct-mapped kernel over every
 * query (no hash form, compressed segments decoded by the kernel).  Every
 * device buffer is carved out of one static array in the library's code
 * object, genie_hip_static_scratch_bytes() bytes per GPU (32 MiB, held by
 * each GPU the library runs on until the process ends); a batch that needs
 * more is refused with GENIE_E_INVALID.  Calls on one GPU share that array:
 * the library runs them one at a time within a process, and the array keeps
 * the last call's contents; calls from several processes on one GPU each
 * have their own copy.  The device address of the array is looked up at
 * genie_hip_init (or at the first call on a device). */
int genie_hip_match_topk_static(genie_index* ix, const genie_dim_t* items, uint64_t n_items, uint32_t n_queries,
                                const uint32_t* k_per_query, uint32_t count_bound, int32_t* out_ids,
                                int32_t* out_counts, void* stream);
/* Size of the static scratch of genie_hip_match_topk_static, bytes. */
int genie_hip_static_scratch_bytes(uint64_t* bytes);

/* ---- codec decode (tests and measurement) ------------------------------ */
/* Decode n_arrays compressed arrays ([cpos[i], cpos[i+1]) words of `words`)
 * on the device with the engine's wave decoders; output i goes to
 * out[i * out_stride ...], its length to out_len[i].  Host buffers. */
int genie_hip_decode(int codec, const uint32_t* words, const uint32_t* cpos, uint32_t n_arrays,
                     uint32_t out_stride, uint32_t* out, uint32_t* out_len);

/* Measurement aid: one device-to-device copy of `bytes` between two fresh
 * buffers (a plain copy kernel of known size, to scale HBM traffic counters). */
int genie_hip_copy_probe(uint64_t bytes);

/* Wait for all work on `stream` (NULL = default stream). */
int genie_hip_stream_synchronize(void* stream);
/* A non-blocking stream of the default priority (a hipStream_t), and its
 * release. */
int genie_hip_stream_create(void** stream);
int genie_hip_stream_destroy(void* stream);
/* Free and total device memory of the current device (cudaMemGetInfo, which
 * knn_MT's admission reads, knn.cu:210). */
int genie_hip_mem_info(uint64_t* free_bytes, uint64_t* total_bytes);
/* Wait for all work on the current device. */
int genie_hip_device_synchronize(void);

const char* genie_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
