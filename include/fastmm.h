/*
 * fastmm.h -- C ABI of the MI355X fast matrix multiplication engine.
 *
 * The reference exposes its hot path as header-only C++ templates, one
 * namespace per algorithm:
 *
 *   template <typename Scalar>
 *   double <alg>::FastMatmul(Matrix<Scalar>& A, Matrix<Scalar>& B, Matrix<Scalar>& C,
 *                            int num_steps, double x = 1e-8,
 *                            Scalar alpha = Scalar(1.0), Scalar beta = Scalar(0.0));
 *   (algorithms/strassen.hpp:662-722; the same shape in every generated header,
 *    e.g. algorithms/fast424_26_206.hpp:2947-3007)
 *
 * with Matrix<Scalar> a column-major view (linalg/linalg.hpp:41-192) and the
 * enum dispatcher RunAlgorithm (algorithms/all_algorithms.hpp:93-243).
 * These entry points replace that path with device pointers, explicit
 * leading dimensions, a status code and no hidden state:
 *
 *   - matrices are column-major; A is m x k, B is k x n, C is m x n;
 *     lda >= max(1, m), ldb >= max(1, k), ldc >= max(1, m); all pointers are
 *     device pointers of the current HIP device;
 *   - C := alpha * A * B + beta * C, computed with num_steps levels of the
 *     named algorithm, dynamic peeling for non-divisible shapes, exactly the
 *     reference's recursion (multipliers folded into the leaf products,
 *     beta applied at the top level only; with num_steps == 0 the reference's
 *     base case ignores beta and so does this one);
 *   - unlike the reference, no caller state changes (the reference leaves
 *     C's multiplier at alpha, algorithms/strassen.hpp:669,380-383).
 *
 * A C++ header with the reference's Matrix<Scalar> / <alg>::FastMatmul
 * spelling over these entry points is include/fastmm/fast_matmul.hpp.
 */
#ifndef FASTMM_H_
#define FASTMM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  FM_OK = 0,
  FM_ERR_INVALID_ARG = 1,   /* null pointer, negative size, ld too small   */
  FM_ERR_UNKNOWN_ALG = 2,   /* name not in the registry                    */
  FM_ERR_SHAPE = 3,         /* num_steps too deep: a block dimension is 0  */
  FM_ERR_WORKSPACE = 4,     /* caller workspace smaller than required      */
  FM_ERR_HIP = 5,           /* HIP runtime error (see fm_last_error)       */
  FM_ERR_UNSUPPORTED = 6    /* dtype/mode combination not available        */
} fm_status;

typedef enum { FM_F64 = 0, FM_F32 = 1 } fm_dtype;

/* Leaf arithmetic.  PLAIN: leaves in the data type on its MFMA path.
 * SPLIT_BF16: FP32 data, leaves from split-bfloat16 products with FP32
 * accumulation.  OZAKI_INT8: FP64 data, leaves from int8 slices with exact
 * int32 accumulation and FP64 reconstruction; the int32 sums stay exact for
 * leaf K (rounded up to 64) up to 2^31 / (S 127^2), i.e. 16640 with the
 * default S = 8 digits: deeper leaves are refused with FM_ERR_SHAPE (use
 * more recursion levels), and the batch autotuner never picks them. */
typedef enum { FM_MODE_PLAIN = 0, FM_MODE_SPLIT_BF16 = 1, FM_MODE_OZAKI_INT8 = 2 } fm_mode;

/* <alg>::FastMatmul (algorithms/strassen.hpp:662-722).  `alg` is the
 * reference's namespace name ("strassen", "grey424_26_206", ...).  `x` is
 * the APA parameter (ignored by exact algorithms).  `stream` is a
 * hipStream_t (NULL = default stream).  `workspace` may be NULL (the library
 * then uses its own cached device arena, one per device and stream, so
 * calls on different streams may run concurrently) or a device buffer of at
 * least fm_workspace_query() bytes.  If `ms_out` is non-NULL the call waits for
 * completion and stores the elapsed device time in ms (the reference's
 * FastMatmul returns its elapsed ms, algorithms/strassen.hpp:708-721). */
fm_status fm_fastmatmul(const char* alg, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                        const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                        int64_t ldc, int num_steps, double x, double alpha, double beta,
                        void* stream, void* workspace, size_t workspace_bytes,
                        float* ms_out);

/* Workspace bytes fm_fastmatmul needs for this problem (the reference sizes
 * its BFS buffers in util/MemoryManager.hpp:65-110). */
fm_status fm_workspace_query(const char* alg, fm_dtype dtype, fm_mode mode, int m, int k,
                             int n, int num_steps, size_t* bytes);

/* ---- per-call options and reusable plans ----------------------------------
 * The reference fixes its schedule at compile time (_PARALLEL_, util/
 * par_util.hpp:17-95) and its leaf kernel by the linked BLAS.  Here every
 * choice a plan embeds is an explicit option, given per call or per plan:
 *   gemm_tile        -1: chosen per leaf launch by the tile cost model;
 *                    0 .. 34 forces an FmTileCfg (see fm_set_gemm_tile);
 *   dfs_levels       top recursion levels run depth-first (0: breadth-first;
 *                    -1: as few as keep the workspace under workspace_limit,
 *                    or, with no limit, under a quarter of the device's
 *                    free memory at plan time);
 *   workspace_limit  bytes (0: none): depth-first levels are added from the
 *                    top until the plan fits (FM_ERR_WORKSPACE otherwise);
 *   split_remainder  1: whole-tile carve-out and the remainder launch;
 *   stage_fusion     1: two-level S/T/W stages where the shape allows, and
 *                    with split_bf16 leaves the last S/T stage writes the
 *                    leaves' bf16 slice planes itself; 2: the same stages,
 *                    the planes cut by a separate pass; 0: one level a time;
 *   ozaki_slices     6, 7 or 8 int8 digits (FM_MODE_OZAKI_INT8).
 * A NULL options pointer means the process defaults, which the fm_set_*
 * calls below change; fm_options_default reads them.  Results are
 * bit-identical under every schedule and tile choice. */
typedef struct {
  int gemm_tile;
  int dfs_levels;
  size_t workspace_limit;
  int split_remainder;
  int stage_fusion;
  int ozaki_slices;
} fm_options;

void fm_options_default(fm_options* out);

/* fm_fastmatmul with explicit options (opts may be NULL). */
fm_status fm_fastmatmul_ex(const char* alg, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                           const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                           int64_t ldc, int num_steps, double x, double alpha, double beta,
                           const fm_options* opts, void* stream, void* workspace,
                           size_t workspace_bytes, float* ms_out);

/* A plan: one multiply's launch sequence, built and uploaded once, run for
 * any A, B, C of the planned shape and leading dimensions.  fm_plan_run
 * issues kernel launches only (no allocation when a workspace is given, no
 * memset), so a caller can capture it into a hipGraph.  No reference
 * counterpart: the reference rebuilds its recursion on every call
 * (algorithms/strassen.hpp:662-722).  alg == NULL plans the classical
 * product (beta honoured). */
typedef struct fm_plan_s* fm_plan;
fm_status fm_plan_create(const char* alg, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                         int64_t lda, int64_t ldb, int64_t ldc, int num_steps, double x,
                         double alpha, double beta, const fm_options* opts, fm_plan* out);
/* workspace == NULL: the library's arena for `stream` (which may allocate
 * on first use: do one run before capturing a graph, or pass a buffer). */
fm_status fm_plan_run(fm_plan plan, const void* A, const void* B, void* C, void* stream,
                      void* workspace, size_t workspace_bytes, float* ms_out);
fm_status fm_plan_workspace(fm_plan plan, size_t* bytes);
void fm_plan_destroy(fm_plan plan);

/* fm_fastmatmul for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * reference counterpart: the reference allocates its temporaries on every
 * call (algorithms/strassen.hpp:414-420).  It calls no hipMalloc, hipFree or
 * allocating library routine: the plan's descriptors and workspace live in a
 * fixed device array of FM_NOALLOC_SCRATCH_BYTES that the library holds on
 * every device it runs on until the process ends, and a multiply needing more
 * fails with FM_ERR_WORKSPACE (use fm_fastmatmul with a workspace then, or
 * fewer levels).  Options are the process defaults.  All calls on one device
 * share the array and leave their contents in it: calls that overlap on one
 * device (two streams, or threads without ordering) race and must not be
 * issued. */
#define FM_NOALLOC_SCRATCH_BYTES 268435456u /* 256 MiB */
fm_status fm_fastmatmul_noalloc(const char* alg, fm_dtype dtype, fm_mode mode, int m, int k,
                                int n, const void* A, int64_t lda, const void* B, int64_t ldb,
                                void* C, int64_t ldc, int num_steps, double x, double alpha,
                                double beta, void* stream, float* ms_out);

/* Classical C := alpha * A * B + beta * C on the same leaf kernels
 * (MatMul, linalg/linalg.hpp:261-269), with the same leaf modes. */
fm_status fm_gemm(fm_dtype dtype, fm_mode mode, int m, int k, int n, const void* A,
                  int64_t lda, const void* B, int64_t ldb, void* C, int64_t ldc, double alpha,
                  double beta, void* stream);

/* ---- batches of independent multiplies -------------------------------------
 * Many C_i := alpha_i A_i B_i + beta_i C_i in one launch sequence: every
 * problem's S/T stages, ONE grouped launch of all base-case products (plus
 * one or two for edge strips and remainder tiles), then the W stages and one
 * grouped peeling launch per level.  A problem with
 * alg == NULL (or num_steps < 0) gets its algorithm (from `candidates`,
 * default "strassen", "grey423_20_144" in its six orientations
 * (grey432/243/234/324/342_20_144) and "smirnov333_23_139") and depth
 * (0 .. max_steps, default 2) from the autotuner's cost model; depth 0 is
 * the classical product with beta honoured.  With use_graph the launch
 * sequence is captured once into a hipGraph and fm_batch_run replays it
 * (the matrices' addresses are fixed at creation).  No reference
 * counterpart: the reference multiplies one problem per call. */
typedef struct {
  const char* alg; /* NULL: autotuned */
  int num_steps;   /* < 0: autotuned */
  int m, k, n;
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  void* C;
  int64_t ldc;
  double alpha, beta;
} fm_problem;

typedef struct fm_batch_s* fm_batch;

fm_status fm_batch_create(fm_dtype dtype, fm_mode mode, const fm_problem* problems, int count,
                          const char* const* candidates, int num_candidates, int max_steps,
                          int use_graph, void* workspace, size_t workspace_bytes,
                          fm_batch* out);
fm_status fm_batch_run(fm_batch batch, void* stream, float* ms_out);
/* algorithm ("classical" at depth 0) and depth chosen for problem `index` */
fm_status fm_batch_choice(fm_batch batch, int index, const char** alg, int* num_steps);
size_t fm_batch_workspace_bytes(fm_batch batch);
void fm_batch_destroy(fm_batch batch);

/* The autotuner's estimate (microseconds) for one multiply; alg == NULL or
 * num_steps == 0 prices the classical product.  The cost model's rates (leaf
 * GEMM TFLOP/s against leaf size per dtype and mode, stage bandwidth, launch
 * and per-level costs) come from a text table measured on the GPU by
 * `python -m fastmm.calibrate`: FM_TUNING_FILE, else tuning/mi355x.txt beside
 * the library, else built-in values.  fm_autotune_load replaces the table;
 * fm_autotune_table names the one in use. */
fm_status fm_autotune_estimate(fm_dtype dtype, fm_mode mode, const char* alg, int m, int k, int n,
                               int num_steps, double* us);
fm_status fm_autotune_load(const char* path);
const char* fm_autotune_table(void);

/* Registry (algorithms/all_algorithms.hpp:52-90, 247-326). */
int fm_num_algorithms(void);
const char* fm_algorithm_name(int index);
fm_status fm_algorithm_info(const char* alg, int* m, int* k, int* n, int* rank);

const char* fm_status_string(fm_status s);
/* Text of the last error on the calling thread ("" if none). */
const char* fm_last_error(void);

/* Per-stage timing.  While enabled, every launch fm_fastmatmul/fm_gemm
 * issues is bracketed by HIP events on the caller's stream (the stream the
 * kernels run on); fm_profile_read waits for them, returns the accumulated
 * device time and launch count per stage kind and resets the counters. */
typedef enum {
  FM_STAGE_S = 0,     /* S_r = sum U A      (HBM-bound)                     */
  FM_STAGE_T = 1,     /* T_r = sum V B      (HBM-bound)                     */
  FM_STAGE_W = 2,     /* C_ij = sum W M (+ beta C)  (HBM-bound)             */
  FM_STAGE_LEAF = 3,  /* grouped leaf products (MFMA-bound)                 */
  FM_STAGE_PEEL = 4,  /* dynamic-peeling products                           */
  FM_STAGE_PREP = 5,  /* split-bf16 / int8-digit plane preparation (HBM)    */
  FM_STAGE_COUNT = 6
} fm_stage_kind;

typedef struct {
  double ms[FM_STAGE_COUNT];
  int64_t launches[FM_STAGE_COUNT];
} fm_profile;

void fm_profile_enable(int on);
fm_status fm_profile_read(fm_profile* out);

/* Measured matrix-unit peak of the current device, for roofline figures (no
 * reference counterpart): every SIMD issues `iters` rounds of independent
 * back-to-back MFMAs on register operands; the result is in tera-operations
 * per second (2 per multiply-add) at the clock the chip holds under that
 * load.  kind 0: FP64 (v_mfma_f64_16x16x4f64), 1: FP32 (16x16x4f32),
 * 2: BF16 (16x16x32_bf16, FP32 accumulate), 3: INT8 (32x32x32_i8). */
fm_status fm_probe_mfma_peak(int kind, int iters, double* tera_ops_per_s);

/* Measured HBM rate of a stage kernel's access pattern on the current
 * device, for roofline figures (no reference counterpart): each lane reads
 * one element position of `reads` FP64 slabs of `elems` elements and writes
 * it to `writes` such slabs; the result is (reads + writes) * elems * 8
 * bytes over the kernel's time, in GB/s.  The S, T and W stages read and
 * write whole slabs this way.  Allocates and frees its buffers (at most
 * 32 GB). */
fm_status fm_probe_stream(int reads, int writes, int64_t elems, double* gb_per_s);

/* Process default of fm_options.ozaki_slices.
 * Digit count S of FM_MODE_OZAKI_INT8 leaves (6, 7 or 8; default 8): each
 * FP64 operand row/column becomes S int8 digits of 7 bits under a
 * power-of-two scale, and S(S+1)/2 int8 products are summed. */
fm_status fm_set_ozaki_slices(int slices);

/* Process default of fm_options.gemm_tile: force leaf-GEMM tile configuration
 * `cfg` (0 .. 34, see FmTileCfg in the sources) for calls that pass no
 * options; -1 restores the automatic choice.  Configurations 24 .. 29 and 31 run
 * split_bf16 leaves on prepared slice planes and are refused in other modes
 * (FM_ERR_INVALID_ARG when a plan is built); in ozaki_int8 mode, 1 (64 x 64)
 * selects the 16x16x64 digit kernel, 2 (128 x 64) the 32x32x32 one on 64 x 64
 * tiles and any other value the default one (32x32x32 on 64 x 128 tiles). */
fm_status fm_set_gemm_tile(int cfg);

/* Diagonal scaling around the fast multiply (scaling/scaling.hpp:105-196,
 * scaling/fast_scaling.cpp): copies of A and B are scaled, scaling_steps
 * times, by the reference's OUTSIDE_INSIDE (0), INSIDE_OUTSIDE (1), INSIDE (2)
 * or OUTSIDE (3) rule -- outer: rows of A and columns of B divided by their
 * largest magnitude; inner: column k of A divided and row k of B multiplied by
 * sqrt(max|A[:,k]| / max|B[k,:]|) -- then multiplied with `alg`, and the
 * product's rows and columns multiplied back (PostProcessScaling):
 * C := alpha * diag(r) (A' B') diag(s) + beta * C.  Improves accuracy when
 * rows of A or columns of B differ widely in magnitude.  A row or column of
 * zeros keeps factor 1.  Workspace: the plan's plus copies of A and B (and
 * of the product when beta != 0); workspace == NULL uses the arena. */
fm_status fm_fastmatmul_scaled(const char* alg, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                               const void* A, int64_t lda, const void* B, int64_t ldb, void* C,
                               int64_t ldc, int num_steps, double x, double alpha, double beta,
                               int scaling_type, int scaling_steps, void* stream, void* workspace,
                               size_t workspace_bytes, float* ms_out);

/* Workspace bytes fm_fastmatmul_scaled needs (its plan's, the copies of A and
 * B, the product when beta != 0, the scaling vectors). */
fm_status fm_scaled_workspace_query(const char* alg, fm_dtype dtype, fm_mode mode, int m, int k,
                                    int n, int num_steps, int beta_nonzero, size_t* bytes);

/* Shape of the plan fm_fastmatmul would run (built on the host, nothing
 * launched; contiguous operands): launches per fm_stage_kind, algorithmic
 * HBM bytes per fm_stage_kind (every source element read once, every output
 * written once), how many of the S/T launches are two-level stages, and the
 * first two leaf launches' tile configuration (-1: int8-digit leaves),
 * product and tile counts. */
typedef struct fm_plan_info {
  int launches[5];       /* FM_STAGE_S .. FM_STAGE_PEEL (preparation counts with LEAF) */
  double stage_bytes[5];
  int two_level_stages;
  int leaf_cfg[2];
  int leaf_products[2];
  int leaf_tiles[2];
  size_t workspace_bytes;
  int dfs_levels;    /* depth-first top levels the plan runs */
  double prep_bytes; /* split-bf16 / int8-digit plane preparation: operands read, planes written */
} fm_plan_info;
fm_status fm_plan_query(const char* alg, fm_dtype dtype, fm_mode mode, int m, int k, int n,
                        int num_steps, fm_plan_info* info);

/* Process defaults of fm_options' schedule fields: the run-time form of the
 * reference's compile-time _PARALLEL_ choice of DFS, BFS or HYBRID
 * (util/par_util.hpp:17-95, algorithms/strassen.hpp:404-463); calls that
 * pass no options follow them.
 *  - dfs_levels = 0 (default): breadth-first, every level one launch per
 *    stage over all of its nodes and one grouped launch of all leaves.
 *  - dfs_levels = d > 0: the top d levels run depth-first, one child at a
 *    time (its whole breadth-first sub-plan), the children reusing one
 *    workspace region: less workspace, smaller leaf launches.
 *  - dfs_levels = -1: the fewest depth-first levels that keep the workspace
 *    under workspace_limit (or a quarter of the free device memory).
 *  - workspace_limit > 0: depth-first levels are added from the top until
 *    the plan's workspace fits the limit (FM_ERR_WORKSPACE if even a fully
 *    depth-first plan does not).
 *  - split_remainder != 0 (default): when the leaf tiles leave the last
 *    round of work groups part-empty, those products run in a second
 *    launch with smaller tiles (the HYBRID rule of running the leftover
 *    leaves with all threads).
 * Results are bit-identical under every schedule. */
fm_status fm_set_schedule(int dfs_levels, size_t workspace_limit, int split_remainder);

/* Process default of fm_options.stage_fusion.
 * Two-level S/T stages (default on): where two consecutive recursion levels
 * have no peeling between them, one launch forms the deeper level's S (T)
 * operands straight from the upper level's, without storing the middle
 * level.  With FM_MODE_SPLIT_BF16 leaves (FP32) whose rows are multiples of
 * 256 and K of 32, the last such stage also writes the leaf operands as
 * bf16 slice planes, which the leaf kernels read, instead of FP32 blocks
 * that a preparation pass would cut into planes.  on: 1 (default), 2
 * (two-level stages, planes by the separate pass), 0 (one level at a
 * time).  Results are bit-identical under all three (plans built afterwards
 * follow the setting).  No counterpart in the reference, whose adds run one
 * level at a time (algorithms/strassen.hpp:9-252). */
void fm_set_stage_fusion(int on);

/* Drop cached plans and the library-owned workspace arenas.  The arenas (one
 * per device and stream) are bounded together by half the device's memory;
 * fm_release_workspace frees the one of `stream` on the current device
 * (after waiting for the device). */
void fm_release_caches(void);
void fm_release_workspace(void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FASTMM_H_ */
