/*
 * mgpu_hip.h -- C ABI of the MI355X (gfx950) sort / merge / scan primitives.
 *
 * Every entry point keeps the reference's two-phase auxiliary-memory protocol
 * (inc/mgpu/vk/mergesort.hxx:35-41, inc/mgpu/vk/merge.hxx:30-34,
 * inc/mgpu/vk/scan.hxx:54-58, inc/mgpu/vk/radix.hxx:288-302):
 *
 *   1. Call with aux == NULL: the required byte count is ADDED to *aux_size
 *      (accumulating, so one pool can be sized for several calls) and nothing
 *      is launched.
 *   2. Call again with aux != NULL and *aux_size == capacity of aux: the work
 *      is enqueued on `stream` (asynchronous, nothing is read back to the
 *      host).  A capacity below the requirement returns MGPU_AUX_TOO_SMALL
 *      (the reference never checked; see inc/mgpu/vk/mergesort.hxx:43-46).
 *
 * As in the reference, a call that needs zero aux bytes still returns early
 * when aux is NULL, so callers pass any non-NULL pointer (a 1-byte buffer)
 * for the second call.  count == 0 is a defined no-op (the reference's
 * find_log2(0) is undefined, inc/mgpu/common/meta.hxx:37-41).
 *
 * All pointers are device pointers (hipMalloc / torch tensors).  Vectorised
 * 128-bit paths are taken when data pointers are 16-byte aligned, otherwise a
 * scalar path runs; results are identical.  aux itself must be 256-byte
 * aligned, as hipMalloc returns it, whenever the call needs any aux bytes
 * (MGPU_INVALID_ARG otherwise); every requirement is a multiple of 256
 * bytes, so consecutive regions of one pool sized by accumulation stay
 * aligned.  Threading: calls on different streams are independent and may
 * come from different host threads; the only global state (the radix
 * algorithm setting, an atomic; the segmented-sort tuning table and the
 * serving engine, each under a mutex) is safe to share.  No C++ exception
 * crosses this boundary.
 */
#ifndef MGPU_HIP_H
#define MGPU_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t without pulling in the HIP headers (same opaque type). */
typedef struct ihipStream_t* mgpu_stream_t;

/* Every enum below spans the whole int32 range (the *_ENUM_MIN / *_ENUM_MAX
 * enumerators name no value): any int a C or FFI caller passes is then a
 * valid value of the type, and one outside the named values reaches the
 * library's argument check (MGPU_INVALID_ARG) instead of being undefined in
 * the C++ that implements the entry points.                                */
#define MGPU_ENUM_RANGE(prefix) prefix##_ENUM_MIN = -0x7fffffff - 1, prefix##_ENUM_MAX = 0x7fffffff

typedef enum mgpu_status_t {
  MGPU_OK = 0,
  MGPU_INVALID_ARG = 1,   /* negative count, bad enum, aux not 256-byte
                             aligned, or a NULL data pointer with count > 0
                             on the call that enqueues (the sizing call with
                             aux == NULL never reads data pointers)          */
  MGPU_AUX_TOO_SMALL = 2, /* *aux_size (capacity) below the requirement          */
  MGPU_HIP_ERROR = 3,     /* a HIP runtime call failed; see mgpu_last_hip_error  */
  MGPU_UNSUPPORTED = 4,   /* size beyond a documented limit of the entry point   */
  MGPU_ENUM_RANGE(MGPU_STATUS)
} mgpu_status_t;

/* Comparator for mergesort / merge (inc/mgpu/vk/mergesort.hxx:11 takes any
 * functor; the C ABI exports the two standard ones, custom functors use the
 * C++ templates in mgpu/hip/ headers). */
typedef enum mgpu_comp_t { MGPU_LESS = 0, MGPU_GREATER = 1, MGPU_ENUM_RANGE(MGPU_COMP) } mgpu_comp_t;

typedef enum mgpu_scan_type_t {
  MGPU_SCAN_EXCLUSIVE = 0, MGPU_SCAN_INCLUSIVE = 1, MGPU_ENUM_RANGE(MGPU_SCAN_TYPE)
} mgpu_scan_type_t;

const char* mgpu_status_string(mgpu_status_t status);
/* hipError_t of the last MGPU_HIP_ERROR raised on the calling thread. */
int mgpu_last_hip_error(void);
/* Library version: major*10000 + minor*100 + patch. */
int mgpu_version(void);
/* Creates the library's per-device state on the current device (the
 * look-back fallback counter: 256 bytes, allocated and zeroed once) and
 * looks up the device address of mgpu_scan_i32_noalloc's scratch array.
 * Optional: any entry point creates what it needs on first use, except
 * while its stream is being captured into a graph, when it runs without the
 * counter (only the fallback count is lost).  Stands in for the reference's
 * context_t setup (src/vk/context.cxx:20-130), of which nothing else is
 * needed here.                                                             */
mgpu_status_t mgpu_init(void);

/* ------------------------------------------------------------------------ */
/* Mergesort: stable, in place.  Replaces vk::mergesort_keys / mergesort_kv /
 * mergesort_indices (inc/mgpu/vk/mergesort.hxx:86-93, 10-84, 95-103).
 * aux: ping-pong keys (+ values), merge-path partitions and, for four-way
 * passes, one sample per tile.
 * Key types: u32, i32, u64, f32, i64, f64.  Values: u32.  indices:
 * vals[i] := i before the sort (inc/mgpu/common/kernel_mergesort.hxx:43-49).
 * count < 2^31.                                                             */
#define MGPU_DECL_MERGESORT(K, KT)                                             \
  mgpu_status_t mgpu_mergesort_keys_##K(void* aux, size_t* aux_size,           \
      mgpu_stream_t stream, KT* keys, int64_t count, mgpu_comp_t comp);        \
  mgpu_status_t mgpu_mergesort_kv_##K##_u32(void* aux, size_t* aux_size,       \
      mgpu_stream_t stream, KT* keys, uint32_t* vals, int64_t count,           \
      mgpu_comp_t comp);                                                       \
  mgpu_status_t mgpu_mergesort_indices_##K(void* aux, size_t* aux_size,        \
      mgpu_stream_t stream, KT* keys, int32_t* indices, int64_t count,         \
      mgpu_comp_t comp);
MGPU_DECL_MERGESORT(u32, uint32_t)
MGPU_DECL_MERGESORT(i32, int32_t)
MGPU_DECL_MERGESORT(u64, uint64_t)
MGPU_DECL_MERGESORT(f32, float)
MGPU_DECL_MERGESORT(i64, int64_t)
MGPU_DECL_MERGESORT(f64, double)
/* Merge pass structure (process-wide): 0 = automatic, 2 = two-way passes
 * only (the reference's structure: one merge-path pass per level), 4 =
 * four-way passes (two levels per pass; when whole tiles leave an odd number
 * of levels the blocksort sorts half tiles, so every pass is four-way) at
 * every size.  Automatic takes four-way passes once a pass moves 168 MiB of
 * keys and values, or 256 MiB of keys only (2^28 u32 pairs: 8 global passes
 * instead of 15).  Results are identical; the sizing call returns enough aux
 * for either.  No counterpart in the reference.                             */
void mgpu_set_mergesort_passes(int ways);
int mgpu_mergesort_passes(void);

/* ------------------------------------------------------------------------ */
/* Two-way stable merge (ties take A first), out of place into c.
 * Replaces vk::merge key-value and key-only (inc/mgpu/vk/merge.hxx:12-46,
 * 49-59) together with vk::merge_path_partitions (inc/mgpu/vk/partition.hxx:
 * 9-20).  aux: one int64 partition per output tile.                         */
#define MGPU_DECL_MERGE(K, KT)                                                 \
  mgpu_status_t mgpu_merge_keys_##K(void* aux, size_t* aux_size,               \
      mgpu_stream_t stream, const KT* a_keys, int64_t a_count,                 \
      const KT* b_keys, int64_t b_count, KT* c_keys, mgpu_comp_t comp);        \
  mgpu_status_t mgpu_merge_kv_##K##_u32(void* aux, size_t* aux_size,           \
      mgpu_stream_t stream, const KT* a_keys, const uint32_t* a_vals,          \
      int64_t a_count, const KT* b_keys, const uint32_t* b_vals,               \
      int64_t b_count, KT* c_keys, uint32_t* c_vals, mgpu_comp_t comp);
MGPU_DECL_MERGE(u32, uint32_t)
MGPU_DECL_MERGE(i32, int32_t)
MGPU_DECL_MERGE(u64, uint64_t)
MGPU_DECL_MERGE(f32, float)
MGPU_DECL_MERGE(i64, int64_t)
MGPU_DECL_MERGE(f64, double)

/* Merge-path partitions, vk::merge_path_partitions<bounds>(cmd, a, a_count,
 * b, b_count, mp_data, spacing, comp) (inc/mgpu/vk/partition.hxx:9-20) with
 * the stream in place of the command buffer and the bounds as an argument:
 * mp[i] = the number of A elements among the first
 * min(i * spacing, a_count + b_count) outputs of the merge of sorted a and b
 * (merge_path, inc/mgpu/common/cta_merge.hxx:10-31), for
 * i <= (a_count + b_count + spacing - 1) / spacing, i.e. the reference's
 * num_merge_partitions (cta_merge.hxx:6-8) entries of int64.  LOWER: ties
 * take A first (the stable merge); UPPER: ties take B first.  No aux; the
 * caller sizes mp.  spacing > 0.                                          */
typedef enum mgpu_bounds_t {
  MGPU_BOUNDS_LOWER = 0, MGPU_BOUNDS_UPPER = 1, MGPU_ENUM_RANGE(MGPU_BOUNDS)
} mgpu_bounds_t;
#define MGPU_DECL_PARTITIONS(K, KT)                                            \
  mgpu_status_t mgpu_merge_path_partitions_##K(mgpu_stream_t stream,         \
      const KT* a, int64_t a_count, const KT* b, int64_t b_count, int64_t* mp, \
      int64_t spacing, mgpu_comp_t comp, mgpu_bounds_t bounds);
MGPU_DECL_PARTITIONS(u32, uint32_t)
MGPU_DECL_PARTITIONS(i32, int32_t)
MGPU_DECL_PARTITIONS(u64, uint64_t)
MGPU_DECL_PARTITIONS(f32, float)
MGPU_DECL_PARTITIONS(i64, int64_t)
MGPU_DECL_PARTITIONS(f64, double)

/* ------------------------------------------------------------------------ */
/* LSD radix sort, 8-bit digits, ascending in radix-bit order (unsigned:
 * value order; signed: MSB flipped; float: -0.0 before +0.0, NaNs last),
 * stable, in place.  Replaces vk::radix_sort (inc/mgpu/vk/radix.hxx:195-470)
 * and extends it with values.  Bit order per inc/mgpu/common/cta_radix.hxx:
 * 6-64.  count < 2^32 - 12288 (32-bit positions, less one tile).  The tile
 * shape is chosen by key size, values and count (small sorts use small
 * tiles); aux sizing follows the same choice.                              */
#define MGPU_DECL_RADIX(K, KT)                                                 \
  mgpu_status_t mgpu_radix_sort_keys_##K(void* aux, size_t* aux_size,          \
      mgpu_stream_t stream, KT* keys, int64_t count);                          \
  mgpu_status_t mgpu_radix_sort_kv_##K##_u32(void* aux, size_t* aux_size,      \
      mgpu_stream_t stream, KT* keys, uint32_t* vals, int64_t count);
MGPU_DECL_RADIX(u32, uint32_t)
MGPU_DECL_RADIX(i32, int32_t)
MGPU_DECL_RADIX(u64, uint64_t)
MGPU_DECL_RADIX(f32, float)
MGPU_DECL_RADIX(i64, int64_t)
MGPU_DECL_RADIX(f64, double)

/* ------------------------------------------------------------------------ */
/* Scan with plus.
 * mgpu_scan_exclusive_T: in place, data[i] := init + sum(data[0..i)) -- the
 *   exact counterpart of vk::scan (inc/mgpu/vk/scan.hxx:10-115), with init
 *   applied once (std::exclusive_scan semantics).
 * mgpu_scan_T: out of place or in place (input == output), exclusive or
 *   inclusive.
 * Integer types wrap modulo 2^32 (two's complement).  f32: float32 sums
 *   inside aligned 256-element blocks (at most 10 additions on any path),
 *   float64 for every sum above a block and for the carries between tiles.
 * mgpu_reduce_T: *out (device) := init + sum(input[0..count)); f32 sums in
 *   float64 and rounds once.                                                */
#define MGPU_DECL_SCAN(T, TT)                                                  \
  mgpu_status_t mgpu_scan_exclusive_##T(void* aux, size_t* aux_size,           \
      mgpu_stream_t stream, TT* data, int64_t count, TT init);                 \
  mgpu_status_t mgpu_scan_##T(void* aux, size_t* aux_size,                     \
      mgpu_stream_t stream, const TT* input, TT* output, int64_t count,        \
      mgpu_scan_type_t type, TT init);                                         \
  mgpu_status_t mgpu_reduce_##T(void* aux, size_t* aux_size,                   \
      mgpu_stream_t stream, const TT* input, int64_t count, TT* out, TT init);
MGPU_DECL_SCAN(i32, int32_t)
MGPU_DECL_SCAN(u32, uint32_t)
MGPU_DECL_SCAN(f32, float)
/* u64 (modulo 2^64): single pass with look-back like the 32-bit scans; a
 * tile's 64-bit sum and flag are published as two atomic words, each with
 * the flag beside one 32-bit half (vk/scan.hxx:10-115 counterpart).        */
MGPU_DECL_SCAN(u64, uint64_t)

/* mgpu_scan_i32 without aux and without any allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and bit-identical results as mgpu_scan_i32; the
 * look-back status words live in a fixed 1 MiB device array that every GPU
 * the library runs on holds until the process ends.  count up to
 * 2,684,354,560 (131072 tiles of 20480 elements); beyond that
 * MGPU_UNSUPPORTED, before any launch.  Calls that overlap on one GPU share
 * that array and race: serialise them (one stream, or events).  The array's
 * address is looked up once per device by mgpu_init or by the first call.   */
mgpu_status_t mgpu_scan_i32_noalloc(mgpu_stream_t stream, const int32_t* input,
                                    int32_t* output, int64_t count,
                                    mgpu_scan_type_t type, int32_t init);

/* Segmented scan (no counterpart in the reference, whose scan has no
 * segmented form): one scan per segment, each starting from init.
 * segments[i] (device, int32, ascending) is the first element of segment i;
 * element 0 always starts a segment; starts at or beyond count are empty.
 *   exclusive: out[i] = init + in[h] + ... + in[i-1]   (h = its segment start)
 *   inclusive: out[i] = init + in[h] + ... + in[i]
 * Single pass: a tile holding a segment start needs no carry beyond it.
 * f32: float32 sums only inside aligned 256-element blocks, float64 above.
 * input == output is allowed.  count < 2^31.                               */
#define MGPU_DECL_SEGSCAN(T, TT)                                               \
  mgpu_status_t mgpu_segmented_scan_##T(void* aux, size_t* aux_size,           \
      mgpu_stream_t stream, const TT* input, TT* output, int64_t count,        \
      const int32_t* segments, int64_t num_segments, mgpu_scan_type_t type,    \
      TT init);
MGPU_DECL_SEGSCAN(i32, int32_t)
MGPU_DECL_SEGSCAN(u32, uint32_t)
MGPU_DECL_SEGSCAN(f32, float)

/* ------------------------------------------------------------------------ */
/* Batched segmented sort (no counterpart in the reference, which has no
 * segmented primitives; the call shape follows the mgpu family's
 * segmented_sort): one stable sort per segment of keys[0..count), where
 * segments[i] (device, int32, ascending, segments[0] = 0) is the first
 * element of segment i and the last segment ends at count.  Segments are
 * bucketed by length and sorted by persistent blocksort kernels (shape per
 * bucket from the tuning table) or, beyond the largest bucket, by a
 * segmented LSD radix sort.  No host synchronisation: the call can be
 * captured in a graph.  count < 2^31.
 * Tuning: `mgpu_segsort_tuning_t` as returned by mgpu_segmented_sort_autotune
 * (measured on this device) or NULL for the library's current table (the
 * default until an autotune of the same key/value/comparator runs).          */
#define MGPU_SEGSORT_MAX_BUCKETS 13
typedef struct mgpu_segsort_tuning_t {
  int num_buckets;
  int cap[MGPU_SEGSORT_MAX_BUCKETS];   /* bucket b holds lengths <= cap[b]    */
  int shape[MGPU_SEGSORT_MAX_BUCKETS]; /* blocksort shape index of bucket b    */
} mgpu_segsort_tuning_t;
/* Blocksort shapes: index -> (nt, vt); returns the number of shapes.      */
int mgpu_segsort_shapes(int* nt, int* vt, int capacity);
#define MGPU_DECL_SEGSORT(K, KT)                                               \
  mgpu_status_t mgpu_segmented_sort_keys_##K(void* aux, size_t* aux_size,      \
      mgpu_stream_t stream, KT* keys, int64_t count, const int32_t* segments,  \
      int64_t num_segments, mgpu_comp_t comp, const mgpu_segsort_tuning_t* t); \
  mgpu_status_t mgpu_segmented_sort_kv_##K##_u32(void* aux, size_t* aux_size,  \
      mgpu_stream_t stream, KT* keys, uint32_t* vals, int64_t count,           \
      const int32_t* segments, int64_t num_segments, mgpu_comp_t comp,         \
      const mgpu_segsort_tuning_t* t);                                          \
  mgpu_status_t mgpu_segmented_sort_autotune_##K(mgpu_stream_t stream,          \
      int has_values, mgpu_comp_t comp, mgpu_segsort_tuning_t* out,            \
      double* ms_table);
MGPU_DECL_SEGSORT(u32, uint32_t)
MGPU_DECL_SEGSORT(i32, int32_t)
MGPU_DECL_SEGSORT(u64, uint64_t)
MGPU_DECL_SEGSORT(f32, float)
MGPU_DECL_SEGSORT(i64, int64_t)
MGPU_DECL_SEGSORT(f64, double)

/* ------------------------------------------------------------------------ */
/* Graph capture of any sequence of library calls on one stream (no
 * counterpart in the reference, which records command buffers,
 * inc/mgpu/vk/context.hxx:93-108).  Between begin and end, calls on `stream`
 * are recorded, not run; mgpu_graph_launch replays them with the same
 * arguments and buffers.  The library clears its look-back state with
 * kernels, never hipMemsetAsync, so replays are exact.                      */
typedef struct mgpu_graph_st* mgpu_graph_t;
mgpu_status_t mgpu_capture_begin(mgpu_stream_t stream);
mgpu_status_t mgpu_capture_end(mgpu_stream_t stream, mgpu_graph_t* graph);
mgpu_status_t mgpu_graph_launch(mgpu_graph_t graph, mgpu_stream_t stream);
void mgpu_graph_destroy(mgpu_graph_t graph);

/* ------------------------------------------------------------------------ */
/* Serving engine (no counterpart in the reference): continuous batching of
 * sort / merge / scan requests over `slots` streams.  A request is admitted
 * as soon as a slot is free and its device bytes fit the HBM budget (first
 * fit from the head of the queue).  It either brings its own device buffers
 * (mgpu_server_submit_buffers) or, as a load generator (mgpu_server_submit),
 * has its input generated on the device from (kind per key type, seed); a
 * generated request's result (and, with keep_inputs, its input) stays
 * device-resident until mgpu_server_release.  Generated requests:
 *   SORT   stable key-value LSD radix sort of n keys, values = index;
 *          keys uniform (f32: finite, U[-1, 1))
 *   MERGE  stable key-value merge of two ascending arrays of n keys each
 *          (values 0..2n-1 in A then B order)
 *   SCAN   u32/i32/u64: exclusive (mod 2^bits) of uniform keys;
 *          f32: inclusive of U[0, 1) (multiples of 2^-24)
 * Timestamps are host steady-clock milliseconds since creation.            */
typedef struct mgpu_server_st* mgpu_server_t;
typedef enum mgpu_req_op_t {
  MGPU_REQ_SORT = 0, MGPU_REQ_MERGE = 1, MGPU_REQ_SCAN = 2, MGPU_ENUM_RANGE(MGPU_REQ_OP)
} mgpu_req_op_t;
typedef enum mgpu_key_type_t {
  MGPU_KEY_U32 = 0, MGPU_KEY_I32 = 1, MGPU_KEY_U64 = 2, MGPU_KEY_F32 = 3,
  MGPU_ENUM_RANGE(MGPU_KEY_TYPE)
} mgpu_key_type_t;
typedef struct mgpu_request_t {
  int32_t op;        /* mgpu_req_op_t   */
  int32_t key_type;  /* mgpu_key_type_t */
  int64_t n;         /* elements (merge: of each input) */
  uint64_t seed;
} mgpu_request_t;
typedef struct mgpu_response_t {
  int64_t id;
  int32_t status;    /* mgpu_status_t of the enqueue */
  int32_t op, key_type, slot;
  int64_t n;
  void* keys;        /* device: sorted keys, merged keys or scan output */
  void* vals;        /* device: u32 values (sort, merge) or NULL          */
  void* input_a;     /* device: input keys (merge: A), NULL unless kept   */
  void* input_b;     /* device: merge B keys, NULL unless kept            */
  double t_submit_ms, t_start_ms, t_done_ms;
} mgpu_response_t;
mgpu_status_t mgpu_server_create(int slots, size_t hbm_budget, int keep_inputs,
                                 mgpu_server_t* out);
double mgpu_server_now(mgpu_server_t server);
/* Device bytes the server budgets for a request.                          */
size_t mgpu_server_request_bytes(const mgpu_request_t* req);
mgpu_status_t mgpu_server_submit(mgpu_server_t server, const mgpu_request_t* req, int64_t* id);
/* A request on the caller's own device buffers (req->seed is unused): the
 * server allocates only the call's aux (which alone counts against the
 * budget), enqueues the work on a slot's stream when admitted, and reports
 * the caller's pointers in the response.  The caller keeps every buffer
 * alive and unmodified until the response is polled.
 *   SORT   in_a (n keys) -> out, sorted stably (out may be in_a).  With
 *          vals_out: values travel with the keys, from vals_a (may equal
 *          vals_out) or, if vals_a is NULL, the original index.
 *   MERGE  in_a (n keys) and in_b (n_b keys, n if n_b == 0), ascending ->
 *          out (n + n_b keys); with vals_out, vals_a and vals_b are merged
 *          alongside (ties take A first).
 *   SCAN   in_a (n) -> out (may be in_a), scan_type exclusive or inclusive,
 *          from 0; integers modulo 2^bits, f32 within the scan's 2e-6 bound. */
typedef struct mgpu_request_buffers_t {
  const void* in_a;
  const void* in_b;
  const uint32_t* vals_a;
  const uint32_t* vals_b;
  void* out;
  uint32_t* vals_out;
  int64_t n_b;
  int32_t scan_type;  /* mgpu_scan_type_t */
  int32_t reserved;
} mgpu_request_buffers_t;
mgpu_status_t mgpu_server_submit_buffers(mgpu_server_t server, const mgpu_request_t* req,
                                         const mgpu_request_buffers_t* bufs, int64_t* id);
/* Completed responses not yet reported (non-blocking); returns how many.  */
int mgpu_server_poll(mgpu_server_t server, mgpu_response_t* out, int max);
/* Blocks until every submitted request has completed (MGPU_OK), or until
 * nothing is in flight and no pending request fits the budget that
 * unreleased results hold (MGPU_AUX_TOO_SMALL: release, then drain again).
 * mgpu_server_submit returns MGPU_UNSUPPORTED for a request larger than the
 * whole budget or with n >= 2^30.                                          */
mgpu_status_t mgpu_server_drain(mgpu_server_t server);
/* Frees a reported response's device buffers and returns its bytes to the
 * budget.  The caller's own reads of those buffers must have completed.    */
mgpu_status_t mgpu_server_release(mgpu_server_t server, int64_t id);
void mgpu_server_stats(mgpu_server_t server, size_t* in_use, size_t* peak, int64_t* pending,
                       int64_t* in_flight);
/* Waits for the requests in flight, drops those not yet started, frees all
 * device buffers the server holds.                                        */
void mgpu_server_destroy(mgpu_server_t server);

/* ------------------------------------------------------------------------ */
/* Synthetic inputs (no counterpart in the reference, whose examples fill
 * host buffers with rand(), e.g. examples/vk_radix2/vk_radix2.cxx:25-27).
 * Counter-based SplitMix64: element i of `kind` for `seed`, written for
 * i in [0, count) as element (offset + i) of the stream.  elem_bytes is 4 or
 * 8 (kind MGPU_GEN_U64 needs 8).                                           */
typedef enum mgpu_gen_kind_t {
  MGPU_GEN_U32 = 0,        /* uniform over [0, 2^32)                         */
  MGPU_GEN_LOWENTROPY = 1, /* AND of four uniform u32 draws                  */
  MGPU_GEN_I32_0_16 = 2,   /* uniform int32 in [0, 16)                       */
  MGPU_GEN_F32_01 = 3,     /* float32 multiples of 2^-24 in [0, 1)           */
  MGPU_GEN_U64 = 4,        /* uniform over [0, 2^64)                         */
  MGPU_GEN_F32_PM1 = 5,    /* float32 multiples of 2^-23 in [-1, 1)          */
  MGPU_GEN_IOTA = 6,       /* element i = i                                  */
  MGPU_GEN_SORTED_U32 = 7, /* ascending: i*s + rand % s, s = 2^32 / count    */
  MGPU_GEN_SORTED_U64 = 8, /* ascending: i*s + rand % s, s = 2^64 / count    */
  MGPU_GEN_SORTED_F32 = 9, /* ascending: (i + U[0,1)) / count                */
  MGPU_ENUM_RANGE(MGPU_GEN)
} mgpu_gen_kind_t;
mgpu_status_t mgpu_generate(mgpu_stream_t stream, void* out, int64_t count, int elem_bytes,
                            mgpu_gen_kind_t kind, uint64_t seed, uint64_t offset);
/* Device-to-device copy of `bytes` (dst and src 16-byte aligned, not
 * overlapping): a plain HBM copy kernel, the practical read + write ceiling
 * that the memory-bound primitives are compared with.  No counterpart in the
 * reference.                                                               */
mgpu_status_t mgpu_copy(mgpu_stream_t stream, void* dst, const void* src, int64_t bytes);

/* ------------------------------------------------------------------------ */
/* Look-back control (no counterpart in the reference, which has no single-
 * pass kernels).  Scan and radix tiles wait for a predecessor's published
 * word at most `limit` polls, then compute that predecessor's contribution
 * from its input, so progress never depends on workgroup dispatch order.
 * Tests set a tiny limit to force that path.  Integer results are
 * identical either way; a float32 scan sums a fallback tile element by
 * element in float64, so its results may differ in the last bits, within
 * the 2e-6 bound stated for the float32 scans.
 * mgpu_lookback_fallbacks returns how often it was taken on the current
 * device (synchronising the device) and optionally resets the count.       */
/* Digit-pass structure of the radix sort (process-wide): 0 = automatic
 * (count + scan + scatter, the faster structure on MI355X at every size),
 * 1 = per-pass count + scan + scatter, 2 = single pass (all-pass histogram,
 * then one look-back kernel per pass) wherever count < 2^30.  Results are
 * identical; the sizing call returns enough aux for either.  No counterpart
 * in the reference (vk/radix.hxx has one structure).                        */
void mgpu_set_radix_algorithm(int algorithm);
int mgpu_radix_algorithm(void);
void mgpu_set_lookback_spin_limit(uint32_t limit);
uint32_t mgpu_lookback_spin_limit(void);
int64_t mgpu_lookback_fallbacks(int reset);

/* ------------------------------------------------------------------------ */
/* Kernel timing (HIP events recorded on the launch stream around every
 * kernel the library enqueues while enabled).  Used by benchmarks to get a
 * per-kernel average duration.  Not thread-safe while enabled.            */
void mgpu_timing_enable(int enable);
/* Synchronises the recorded events and returns the number of distinct kernel
 * names; for i < capacity fills names[i] (pointer valid until the next
 * mgpu_timing_enable call), launches[i] and total_ms[i].                   */
int mgpu_timing_collect(const char** names, int* launches, double* total_ms,
                        int capacity);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* MGPU_HIP_H */
