/*
 * stablebear_amd: MI355X-native (gfx950) pairwise PCF integration.
 *
 * C-ABI drop-in for the reference's GPU plug-in point. Plain C types only; no
 * C++ types or exceptions cross this boundary. Every function returns an int
 * status (SBAMD_OK on success); on failure sbamd_last_error() gives a
 * thread-local message.
 *
 * Data model. A PCF set is CSR: offsets[n + 1] (int64, host), and the
 * breakpoint times t[offsets[n]] and values v[offsets[n]] in one dtype
 * (SBAMD_F32 for the reference's pcf32, SBAMD_F64 for pcf64). PCF i takes the
 * value v[k] on [t[k], t[k+1]) and v[last] on [t[last], inf). Validation
 * follows the reference's Pcf constructor (src/python/pypcf_support.hpp:36-89):
 * times non-decreasing, first time exactly 0; an empty PCF is the zero PCF
 * [(0, 0)]. Times must also be finite.
 *
 * Output layouts (written by the device directly, 64-bit indices):
 *   SBAMD_OUT_DISTMAT  D[i,j], i != j, at max(i,j)*(max(i,j)-1)/2 + min(i,j)
 *                      (include/sbear/distance_matrix.hpp:120-146)
 *   SBAMD_OUT_SYMMAT   K[i,j] at max(i,j)*(max(i,j)+1)/2 + min(i,j)
 *                      (include/sbear/symmetric_matrix.hpp:80-103)
 *   SBAMD_OUT_DENSE    X[i,j] at i * n_cols + j (py_distance.cpp:156-160)
 */
#ifndef STABLEBEAR_AMD_H
#define STABLEBEAR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define SBAMD_OK 0
#define SBAMD_ERR_INVALID 1   /* bad argument or input (maps to ValueError) */
#define SBAMD_ERR_RUNTIME 2   /* HIP / runtime failure (maps to RuntimeError) */
#define SBAMD_ERR_NODEVICE 3  /* no HIP device */
#define SBAMD_ERR_CANCELLED 4 /* job was cancelled */
#define SBAMD_TIMEOUT 5       /* sbamd_job_wait: not finished yet */
#define SBAMD_ERR_NOMEM 6     /* device or host allocation failed */

/* ---- enums ---- */
#define SBAMD_F32 1
#define SBAMD_F64 2

#define SBAMD_OP_L1 0  /* |f - g|, identity finalize      (operations.cuh:19-31) */
#define SBAMD_OP_LP 1  /* |f - g|^p, finalize x^(1/p)      (operations.cuh:33-51) */
#define SBAMD_OP_DOT 2 /* f * g, identity finalize         (operations.cuh:53-65) */

#define SBAMD_OUT_DISTMAT 0
#define SBAMD_OUT_SYMMAT 1
#define SBAMD_OUT_DENSE 2

/* Arithmetic forms.
 * FP64  : times and values in FP64, one FMA per segment in walk order; the
 *         same operation sequence as the reference CPU integral.
 * FP32C : breakpoint comparisons and segment lengths in the input time dtype,
 *         values and segment products in FP32, Kahan-compensated FP32
 *         accumulation per pair, final per-pair combination in FP64; pairs
 *         whose walk spans more than the plan's fold threshold of segments
 *         fold their partial sums into an FP64 accumulator at every phase. */
#define SBAMD_PREC_FP64 0
#define SBAMD_PREC_FP32C 1

typedef struct sbamd_ctx sbamd_ctx;
typedef struct sbamd_set sbamd_set;
typedef struct sbamd_plan sbamd_plan;
typedef struct sbamd_job sbamd_job;
typedef struct sbamd_graph sbamd_graph;

typedef struct
{
  int64_t n;               /* number of PCFs */
  const int64_t* offsets;  /* n + 1 entries, host memory */
  const void* t;           /* offsets[n] times, host memory */
  const void* v;           /* offsets[n] values, host memory */
  int32_t dtype;           /* SBAMD_F32 or SBAMD_F64 */
} sbamd_pcf_set;

#define SBAMD_HOST 0
#define SBAMD_DEVICE 1

/* Strided view of a CSR PCF set in host or device memory: point k's time is
 * at (char*)t + k * t_stride and its value at (char*)v + k * v_stride, so SoA
 * arrays (stride = element size) and AoS {t, v} points (stride = twice the
 * element size, v = t + element size) are both views. */
typedef struct
{
  int64_t n;               /* number of PCFs */
  const int64_t* offsets;  /* n + 1 entries, in `location` memory */
  const void* t;
  const void* v;
  int64_t t_stride;        /* bytes */
  int64_t v_stride;        /* bytes */
  int32_t dtype;           /* SBAMD_F32 or SBAMD_F64 */
  int32_t location;        /* SBAMD_HOST or SBAMD_DEVICE */
} sbamd_pcf_view;

/* Pair-tile shape of the tile kernel: rows x cols pairs per workgroup
 * (rows + cols <= 64), `cap` breakpoint records per PCF staged in LDS per
 * phase, `lane_pairs` pairs walked by each lane (1 or 2; a workgroup has
 * rows * ceil(cols / lane_pairs) lanes, at most 512). Zero fields mean "use
 * the tuned or default value"; for a pdist / kernel plan, rows without cols
 * means a square tile. Tiles of pdist / kernel plans may be rectangular:
 * those that reach the diagonal skip the pairs on or above it. */
typedef struct
{
  int32_t rows;
  int32_t cols;
  int32_t cap;
  int32_t lane_pairs;
} sbamd_tile_shape;

typedef struct
{
  uint64_t pairs;          /* output entries computed */
  uint64_t segments;       /* total union segments walked (exact, host count) */
  uint64_t out_count;      /* output elements */
  uint64_t tiles;          /* workgroup tiles */
  uint64_t input_bytes;    /* device bytes of the staged record arrays */
  int32_t rows, cols, cap; /* tile shape in use */
  int32_t problems;        /* 1, or k for a batched plan */
  int32_t long_pairs;      /* 1 if the plan runs the long-pair kernel */
  int32_t lane_pairs;      /* pairs per lane of the tile kernel */
} sbamd_plan_stats;

/* ---- status / device ---- */
const char* sbamd_last_error(void);
const char* sbamd_version(void);
int sbamd_device_count(int* count);

/* ---- context (one HIP device, one default stream) ----
 * Replaces the per-device state of CudaBlockPipeline / PcfBlockOp
 * (include/sbear/cuda/cuda_block_executor.cuh:108-341, pcf_block_op.cuh:32-123). */
int sbamd_ctx_create(int device, sbamd_ctx** out);
int sbamd_ctx_destroy(sbamd_ctx* ctx);
int sbamd_ctx_synchronize(sbamd_ctx* ctx);
/* The context's own (non-blocking) stream, for callers without a stream. */
int sbamd_ctx_stream(sbamd_ctx* ctx, void** stream);
/* Device bytes the host-to-host jobs (sbamd_pairwise_submit) may give their
 * output at once; 0 (the default) = half the free device memory when the job
 * starts. Outputs above it, or above 32 MiB, run in bands of rows whose
 * downloads overlap the next bands' compute. Replaces the reference's output
 * sizing from free memory (block_matrix_support.cuh:69-93). */
int sbamd_ctx_set_output_budget(sbamd_ctx* ctx, uint64_t bytes);
/* Tile shape of the jobs' plans (zero fields: defaults). Replaces the
 * reference's block-size settings (stablebear/system.py set_block_size). */
int sbamd_ctx_set_tile_shape(sbamd_ctx* ctx, const sbamd_tile_shape* shape);

/* ---- device-resident PCF sets ----
 * Validates, drops zero-length segments, sorts PCFs by record count and
 * uploads CSR record arrays laid out for `precision` (one H2D per set).
 * Replaces init_pcf_data / OffsetDataManager / upload_subset
 * (cuda_pcf_data_manager.cuh:13-23, offset_data_manager.hpp:42-68,
 * cuda_offset_data_manager.cuh:24-44). */
int sbamd_set_create(sbamd_ctx* ctx, const sbamd_pcf_set* set, int32_t precision, sbamd_set** out);
/* The same set, bit for bit, validated and packed on the GPU (one wavefront
 * per PCF): host views are copied to the device once as they lie, device
 * views are read in place. Same errors and messages as sbamd_set_create.
 * Replaces the same reference code, which flattens every point on one host
 * thread (cuda_pcf_data_manager.cuh:13-23, offset_data_manager.hpp:42-68). */
int sbamd_set_create_gpu(sbamd_ctx* ctx, const sbamd_pcf_view* view, int32_t precision, sbamd_set** out);
int sbamd_set_destroy(sbamd_set* set);
int sbamd_set_info(const sbamd_set* set, int64_t* n, int64_t* records, int64_t* points);
/* Copy the packed device layout to host buffers (any may be NULL): records
 * (records x *rec_bytes bytes, sorted order), sorted record offsets (n + 1),
 * sorted position -> PCF (n), record -> source point (records; -1 for the
 * record of an empty PCF). No reference counterpart: it exposes the device
 * layout for inspection and tests. */
int sbamd_set_export(const sbamd_set* set, void* rec, int64_t* roff, int64_t* perm, int64_t* src,
                     int32_t* rec_bytes);
/* Replace the values of an existing set from a device array laid out like the
 * original v (offsets[n] entries, set dtype); times and structure unchanged. */
int sbamd_set_update_values(sbamd_set* set, const void* v_device, void* stream);
/* L_p norm of every PCF of a device set, (integral over [0, MAX) of |f|^p)^(1/p)
 * with MAX the largest finite value of the set's time type; p >= 1. Writes n
 * values of out_dtype (SBAMD_F32 / SBAMD_F64) to out_device in the original
 * PCF order, asynchronously on `stream` (NULL: the legacy default stream).
 * Replaces l1_norm / lp_norm applied per element by ApplyFunctional
 * (include/sbear/algorithms/functional/apply_functional.hpp:63-100,
 * src/python/functional/py_norms.cpp:24-46). */
int sbamd_set_lp_norm(sbamd_set* set, double p, void* out_device, int32_t out_dtype, void* stream);

/* ---- plans: tile work list for one pairwise operation ----
 * cols == NULL: pdist (DISTMAT) or l2_kernel (SYMMAT) over rows;
 * cols != NULL: cdist (DENSE, n_rows x n_cols). Replaces CudaBlockScheduler +
 * get_max_output_elements / get_min_block_side / get_grid_dims
 * (cuda_block_scheduler.hpp:29-129, block_matrix_support.cuh:26-93). */
int sbamd_plan_create(sbamd_ctx* ctx, const sbamd_set* rows, const sbamd_set* cols, int32_t op,
                      double p, int32_t out_kind, const sbamd_tile_shape* shape, sbamd_plan** out);
/* Batched pdist / l2_kernel over k independent sets in one launch; outputs
 * are concatenated (problem b at sbamd_plan_out_offset(plan, b)). The plan
 * keeps its own copy of the member sets' records; members whose values were
 * updated (sbamd_set_update_values) are copied again, on the launch stream,
 * by every execute and graph launch. The member sets must outlive the plan. */
int sbamd_plan_create_batched(sbamd_ctx* ctx, int32_t k, const sbamd_set* const* sets, int32_t op,
                              double p, int32_t out_kind, const sbamd_tile_shape* shape,
                              sbamd_plan** out);
int sbamd_plan_destroy(sbamd_plan* plan);
int sbamd_plan_info(const sbamd_plan* plan, sbamd_plan_stats* info);
int sbamd_plan_out_offset(const sbamd_plan* plan, int32_t problem, uint64_t* offset);
/* Pairs whose union walk spans at least this many segments fold FP32C
 * partials into FP64 per phase (default 8192). */
int sbamd_plan_set_fold_threshold(sbamd_plan* plan, int64_t segments);
/* Persistent launch (on by default for batched plans): as many workgroups as
 * fit on the device take tiles from atomic counters that a fill kernel resets
 * before each launch (graph-capture safe). A batched plan keeps a work queue
 * per XCD (units of tiles that share their PCFs, dealt largest first to the
 * least-loaded queue, longest tiles first within a unit); a workgroup takes
 * from its XCD's queue, then from the others. Off: one workgroup per tile. */
int sbamd_plan_set_persistent(sbamd_plan* plan, int32_t on);
/* Long-pair kernel: one wavefront per pair, the pair's union segments split
 * across the 64 lanes at merge-path points, partial sums added in FP64. For
 * plans of few, long pairs (a small N of long PCFs, lp_distance of long
 * PCFs), where one lane per pair cannot fill the device. mode 0: never;
 * 1: always (FP64 results then differ from the sequential sum in the last
 * bits); 2 (default): FP32C plans with fewer than ~384 pairs per CU and
 * >= 2048 segments per pair on average. Not for batched plans. No
 * reference counterpart (the reference walks every pair on one thread,
 * cuda_pcf_kernel.cuh:95-136). */
int sbamd_plan_set_long_pairs(sbamd_plan* plan, int32_t mode);

/* Enqueue the whole forward computation on `stream` (a hipStream_t; NULL is
 * the null stream, as in HIP). out_device: device buffer of plan out_count elements of
 * out_dtype. row_sum_device (optional, DENSE only): n_rows FP64 sums of each
 * output row (for row means). Asynchronous; capturable into a hipGraph.
 * Replaces launch_pcf_block_integrate + the block pipeline's scatter
 * (cuda_pcf_kernel.cuh:95-144, cuda_result_writer.hpp:46-108). */
int sbamd_plan_execute(sbamd_plan* plan, void* out_device, int32_t out_dtype, void* row_sum_device,
                       void* stream);

/* Synchronous convenience: run the plan into a library-owned device buffer
 * and copy the result to out_host (plan out_count elements of out_dtype). */
int sbamd_plan_execute_to_host(sbamd_plan* plan, void* out_host, int32_t out_dtype);

/* The plain FP64 GPU kernel: one lane per pair walking the raw breakpoints
 * of the two PCFs in global memory, in the original PCF order, with the
 * reference walk (iterate_rectangles.hpp:37-104) and integral
 * (matrix_integrate.hpp:18-29). The same design as the reference's CUDA
 * kernel (cuda_pcf_kernel.cuh:95-136), kept as the measured GPU baseline
 * beside the tiled kernels. FP64 sets only; out_kind DISTMAT or SYMMAT. */
typedef struct sbamd_plain sbamd_plain;
int sbamd_plain_create(sbamd_ctx* ctx, const sbamd_pcf_set* set, sbamd_plain** out);
int sbamd_plain_execute(sbamd_plain* plain, int32_t op, double p, int32_t out_kind, void* out_device,
                        void* stream);
int sbamd_plain_destroy(sbamd_plain* plain);

/* Backward of L = sum_e grad_out[e] * out[e] over the plan's output entries
 * (no reference counterpart). op must be SBAMD_OP_LP with p == 2 (fwd_out =
 * the forward distances) or SBAMD_OP_DOT. grad_v / grad_t are device arrays
 * indexed like the set's original points (FP32 or FP64 per grad_dtype),
 * overwritten. Single-set plans only. The first backward of a plan sorts
 * the set's breakpoints into the time-ordered event lists the kernel walks
 * and keeps them with the plan's working arrays; later backwards of the plan
 * (a set's times never change) only refresh the lists' value jumps. */
int sbamd_plan_backward(sbamd_plan* plan, const void* grad_out_device, const void* fwd_out_device,
                        int32_t out_dtype, void* grad_v_device, void* grad_t_device,
                        int32_t grad_dtype, void* stream);

/* The same backward with no device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under hipGraph capture,
 * tight training or serving loops). Its working arrays live in one static
 * device array of about 67 MiB per GPU, held until the process ends, sized
 * for sets of at most 65,536 PCFs and 524,288 records (larger sets:
 * SBAMD_ERR_INVALID). The plan itself is made beforehand. Every call on a GPU
 * shares that GPU's array (each clears what it accumulates into): calls
 * that overlap on one GPU, on any streams, race. */
int sbamd_plan_backward_noalloc(sbamd_plan* plan, const void* grad_out_device, const void* fwd_out_device,
                                int32_t out_dtype, void* grad_v_device, void* grad_t_device,
                                int32_t grad_dtype, void* stream);

/* ---- hipGraph capture of a plan's forward launch ----
 * The graph binds out_device, the plan's tile list and launch parameters and
 * the sets' device arrays. sbamd_graph_launch refuses (SBAMD_ERR_INVALID) a
 * graph whose plan was since reshaped (sbamd_plan_set_shape), given a new
 * fold threshold, long-pair mode or persistence, autotuned or destroyed;
 * capture a new graph then. Value updates of the sets need no new capture. */
int sbamd_graph_create(sbamd_plan* plan, void* out_device, int32_t out_dtype, sbamd_graph** out);
int sbamd_graph_launch(sbamd_graph* graph, void* stream);
int sbamd_graph_destroy(sbamd_graph* graph);

/* ---- autotuner ----
 * PCFs fall into four length classes by record count (below 128, 512, 2048,
 * and the rest); the pairs whose longer PCF is of class k form class k's
 * tiles, run in a launch of their own. sbamd_autotune times candidate tile
 * shapes (limited by the LDS budget) on each class's launch and keeps the
 * fastest per class, caching it by device architecture, operation, record
 * layout, plan kind, class and a coarse histogram of the class (PCF count,
 * mean record count, tile count): a cache hit skips the timing (the cache
 * lives for the process, and in the file SBAMD_TUNE_CACHE names, if set).
 * Separate class launches each end with a tail of their longest tiles: the
 * plan keeps the per-class shapes only when their launches beat one
 * longest-first launch of a single shape (its own or a class winner).
 * For FP32C plans (or plans set to long-pair mode 1) the long-pair kernel is
 * a candidate for the whole plan too; when it wins the plan uses it
 * (sbamd_plan_info long_pairs). *best receives the shape of the class with
 * the most tiles and *best_ms the tuned plan's time. No reference
 * counterpart beyond the fixed block sizing of cuda_block_scheduler.hpp:29-129. */
int sbamd_autotune(sbamd_plan* plan, void* out_device, int32_t out_dtype, int32_t reps,
                   sbamd_tile_shape* best, double* best_ms);
/* The tile shape and launch tile count of each of the four length classes
 * (shapes[4], tiles[4]; either may be NULL). When every class runs in one
 * launch its tiles are counted in class 0; the shapes are then the ones the
 * autotuner found per class (or the plan's, if none were tuned or set). */
int sbamd_plan_class_shapes(const sbamd_plan* plan, sbamd_tile_shape* shapes, int64_t* tiles);
/* Give each length class its own tile shape (shapes[4]; zero fields: the
 * plan's); its tiles then run in a launch of their own. */
int sbamd_plan_set_class_shapes(sbamd_plan* plan, const sbamd_tile_shape* shapes);
/* Length classes the autotuner found in its cache (hits) or had to time
 * (misses), since the process started. */
int sbamd_autotune_cache_stats(uint64_t* hits, uint64_t* misses);
int sbamd_plan_set_shape(sbamd_plan* plan, const sbamd_tile_shape* shape);

/* ---- asynchronous host-to-host jobs (the reference's StoppableTask) ----
 * Runs set upload, planning, execution and the download into out_host on a
 * worker thread. Outputs within the context's output budget (and 32 MiB)
 * are computed whole on the device and downloaded once; larger ones in at
 * least 8 bands of consecutive rows (original order), each band's output
 * written in place on the device, downloaded into pinned memory (kept by the
 * context for later jobs) while the next bands compute, and copied into
 * out_host. progress counts pairs ("integral" unit).
 * Replaces create_cuda_block_{integrate_l1,integrate_lp,integrate_l2_kernel,
 * cdist_l1,cdist_lp}_task (cuda_matrix_integrate_api.hpp:19-113) and the
 * StoppableTask wait / progress / cancel protocol (task.hpp:15-193). */
int sbamd_pairwise_submit(sbamd_ctx* ctx, const sbamd_pcf_set* rows, const sbamd_pcf_set* cols,
                          int32_t op, double p, int32_t out_kind, int32_t precision, void* out_host,
                          int32_t out_dtype, sbamd_job** out);
/* The same job from strided views (sbamd_pcf_view: AoS {t, v} points, as a
 * list of (m, 2) point arrays concatenates to, need no split into time and
 * value arrays), host or device. The offsets are checked at submission; the
 * points by the GPU packer during the job, whose wait then returns the first
 * offending PCF's error. Serves the same reference factories as
 * sbamd_pairwise_submit (cuda_matrix_integrate_api.hpp:19-113), whose
 * Python callers hand over lists of PCFs. */
int sbamd_pairwise_submit_view(sbamd_ctx* ctx, const sbamd_pcf_view* rows, const sbamd_pcf_view* cols,
                               int32_t op, double p, int32_t out_kind, int32_t precision, void* out_host,
                               int32_t out_dtype, sbamd_job** out);
/* SBAMD_OK when finished, SBAMD_TIMEOUT if still running after timeout_ms
 * (< 0 waits forever); the job's own error status once finished. */
int sbamd_job_wait(sbamd_job* job, int64_t timeout_ms);
int sbamd_job_progress(sbamd_job* job, uint64_t* done, uint64_t* total);
int sbamd_job_cancel(sbamd_job* job);
int sbamd_job_destroy(sbamd_job* job);

/* ---- single-pair helpers (synchronous) ---- */
/* lp_distance(f, g, p): lp_distance.hpp:10-21. f, g: one-PCF sets. */
int sbamd_lp_distance(sbamd_ctx* ctx, const sbamd_pcf_set* f, const sbamd_pcf_set* g, double p,
                      double* out);
/* lp_norm(fs, p) of stablebear/norms.py:10-53 (backends Norms_f32_f32 /
 * Norms_f64_f64, src/python/functional/py_norms.cpp:24-68): uploads the host
 * set in the layout of `precision`, runs sbamd_set_lp_norm and copies the n
 * norms (out_dtype) to out_host; synchronous. */
int sbamd_lp_norm(sbamd_ctx* ctx, const sbamd_pcf_set* set, double p, int32_t precision,
                  void* out_host, int32_t out_dtype);
/* Rectangles of the walk over [a, b): iterate_rectangles.hpp:37-104. out
 * holds 4 values (left, right, f_value, g_value) per rectangle in the set
 * dtype; *count receives the total (only the first cap are written). */
int sbamd_iterate_rectangles(sbamd_ctx* ctx, const sbamd_pcf_set* f, const sbamd_pcf_set* g,
                             double a, double b, void* out, int64_t cap, int64_t* count);

/* ---- host-only planning helpers (no device needed) ---- */
/* ---- row reductions of a pairwise output in device memory ----
 * `out_kind` SBAMD_OUT_DISTMAT: the compressed lower triangle of an n x n
 * distance matrix (n = n_rows; row i runs over every j != i); SBAMD_OUT_DENSE:
 * a row-major n_rows x n_cols matrix. Elements of `dtype` (SBAMD_F32 /
 * SBAMD_F64); asynchronous on `stream` (NULL: the null stream). No reference
 * counterpart: callers of the reference reduce the returned matrices on the
 * host (D.mean(axis=1) in its cdist benchmark); these generalise the fused
 * row sums of sbamd_plan_execute to any output already on the device.
 * Per row: sum, minimum and maximum in FP64 (any output may be NULL; the
 * mean is sum / entries). */
int sbamd_rows_reduce(sbamd_ctx* ctx, int32_t out_kind, int64_t n_rows, int64_t n_cols, const void* d_device,
                      int32_t dtype, double* sum_device, double* min_device, double* max_device, void* stream);
/* The k nearest columns of every row (1 <= k <= 32): values (`dtype`) and
 * column indices, n_rows x k row-major, ascending by (value, column); rows
 * with fewer than k entries are padded with +inf and -1. */
int sbamd_rows_knn(sbamd_ctx* ctx, int32_t out_kind, int64_t n_rows, int64_t n_cols, const void* d_device,
                   int32_t dtype, int32_t k, void* values_device, int64_t* index_device, void* stream);

/* ---- the reference's binary file format (include/sbear/io.hpp:27-158,
 * io/tensor_io.hpp, io/pcf_io.hpp, io/compressed_matrix_io.hpp) ----
 * Files hold one tensor or one object behind a header ("\1MPCF", 'e', format
 * version, three length-prefixed strings). These functions work on whole
 * file images in memory; format errors return SBAMD_ERR_RUNTIME with the
 * reference's messages. */
#define SBAMD_IO_TENSOR 1
#define SBAMD_IO_OBJECT 2
typedef struct
{
  int32_t version;      /* format version of the file (1 or 2) */
  uint32_t kind;        /* SBAMD_IO_TENSOR or SBAMD_IO_OBJECT */
  int32_t base_format;  /* element type (tensor_io.hpp:92-121): 1 float, 2 int, 3 uint,
                           4 bool, 100 PCF, 101 integer PCF, 1000 point cloud,
                           1100 symmetric matrix, 1120 distance matrix, 10000 barcode */
  int32_t sub_format;   /* element width in bits */
  int32_t ndim;         /* tensor rank (0 for objects) */
  int32_t reserved;
  uint64_t shape[32];
  uint64_t elements;    /* elements in row-major order (a rank-0 tensor has none; an object one) */
  uint64_t payload;     /* byte offset of the first element */
} sbamd_io_info;
/* Parse the header, format and shape of a file image
 * (replaces sb::read_any_tensor / read_any_object, io.hpp:222-300). */
int sbamd_io_parse(const void* buf, uint64_t len, sbamd_io_info* info);
/* Byte extent of every element: for PCFs counts[e] = points and starts[e] =
 * byte offset of the first (time, value) pair; for compressed matrices
 * counts[e] = n and starts[e] = offset of the first stored value; numbers
 * are contiguous (counts 1). *end = offset past the last element. Checks
 * that everything lies inside the image (replaces read_tensor,
 * tensor_io.hpp:213-241). Point clouds and barcodes: SBAMD_ERR_INVALID. */
int sbamd_io_index(const void* buf, uint64_t len, const sbamd_io_info* info, uint64_t* counts,
                   uint64_t* starts, uint64_t* end);
/* Header, kind, format and (for tensors) shape of a file holding one tensor
 * or object; *size = bytes (out may be NULL to size). Replaces write_header +
 * write_contiguous_tensor's prologue (io.hpp:119-127, tensor_io.hpp:166-177). */
int sbamd_io_header(uint32_t kind, int32_t base_format, int32_t sub_format, int32_t ndim,
                    const uint64_t* shape, void* out, uint64_t cap, uint64_t* size);
/* PCFs of a CSR set as file elements, [u64 points][(t, v) x points] each, in
 * order; *size = bytes (out may be NULL to size). Replaces the PCF element
 * writer (pcf_io.hpp:10-14, point_io.hpp:10-15). */
int sbamd_io_write_pcfs(int64_t n, const int64_t* offsets, const void* t, const void* v, int32_t dtype,
                        void* out, uint64_t cap, uint64_t* size);
/* Device set straight from a file image holding a tensor (any shape,
 * row-major order) or an object of pcf32 / pcf64: the PCF payload is copied
 * to the device once as it lies and validated and packed on the GPU. */
int sbamd_set_create_image(sbamd_ctx* ctx, const void* buf, uint64_t len, int32_t precision, sbamd_set** out);

/* Validate a host set like sbamd_set_create would; returns the status. */
int sbamd_validate_set(const sbamd_pcf_set* set);
/* Count union segments of the walk over all pairs of a pdist (cols == NULL,
 * include_diagonal 0), l2_kernel (cols == NULL, include_diagonal 1) or cdist
 * (cols != NULL), exactly, after zero-length segments are dropped. */
int sbamd_count_segments(const sbamd_pcf_set* rows, const sbamd_pcf_set* cols,
                         int32_t include_diagonal, uint64_t* segments);
/* Tile list the planner would build for record counts m_rows (n_rows) and
 * m_cols (n_cols, or NULL for the strict lower triangle): writes up to cap
 * tiles as (row0, col0, nrows, ncols) int32 quadruples in sorted index space
 * plus the sorted permutation of rows (perm_rows, n_rows entries). In the
 * triangle, ncols carries 1 << 16 when the tile also holds pairs on or
 * above the diagonal (the kernel then skips those); tiles may be
 * rectangular there too. */
int sbamd_plan_tiles_host(int64_t n_rows, const int64_t* m_rows, int64_t n_cols,
                          const int64_t* m_cols, int32_t tile_rows, int32_t tile_cols,
                          int32_t* tiles, int64_t cap, int64_t* n_tiles, int64_t* perm_rows);
/* The per-XCD work queues a batched (persistent) plan gives a tile list (no
 * reference counterpart; a planner check): tiles are (row0, col0, problem)
 * int32 triples of tile_rows x tile_cols tiles, in longest-first order, with
 * their work; n_queues queues. order receives the new position of every
 * input tile's index (the reordered list is tiles[order[0]], tiles[order[1]],
 * ...), queue_off the n_queues + 1 queue starts in it. */
int sbamd_queue_order_host(int64_t n_tiles, const int32_t* tiles, const int64_t* work, int32_t tile_rows,
                           int32_t tile_cols, int32_t n_queues, int64_t* order, int64_t* queue_off);

#ifdef __cplusplus
}
#endif

#endif
