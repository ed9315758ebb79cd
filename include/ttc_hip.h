/*
 * ttc_hip.h — C ABI of libttc_hip.so, the MI355X (gfx950) tensor-transposition
 * runtime that replaces HPAC/TTC's generated CUDA path.
 *
 * Operation (column-major, index 0 is stride-1, as in TTC):
 *     B[pi(i)] = alpha * A[i] + beta * B[pi(i)]
 * where B's dimension q is A's dimension perm[q] (ttc/transpose.py:462-473,
 * validity check ttc/ttc.py:739-742).  size[k] is A's extent k; lda[k] is the
 * outer extent of A's dimension k and ldb[q] the outer extent of B's dimension
 * q (NULL = dense), so A's stride of dimension k is prod_{m<k} lda[m]
 * (ttc/transpose.py:443-473).  When beta_is_zero is set, B is not read
 * (ttc/transpose.py:133-136, ttc/CUDAtranspose.py:909-912).
 *
 * alpha is applied in A's real type and beta in B's real type
 * (ttc/transpose.py:402-407, ttc/CUDAtranspose.py:22-32); the API takes them
 * as double and rounds them to those types.
 *
 * Pointers A and B are device pointers owned by the caller.  Host arrays
 * (perm, size, lda, ldb) are copied into the plan; nothing is uploaded per
 * call (unlike the reference wrapper, ttc/CUDAtranspose.py:403-415).
 */
#ifndef TTC_HIP_H
#define TTC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TTC_MAX_DIM 8       /* rank limit after index fusion */
#define TTC_MAX_INPUT_DIM 16 /* rank limit before index fusion */

/* Element types.  S, D, C, Z are TTC's s/d/c/z (ttc/ttc.py:1253-1280,
 * complex = interleaved (re, im) pairs, the cuFloatComplex/cuDoubleComplex
 * layout).  BF16, F16, E4M3, E5M2 have no counterpart in the reference. */
typedef enum {
  TTC_S = 0,    /* float */
  TTC_D = 1,    /* double */
  TTC_C = 2,    /* float complex */
  TTC_Z = 3,    /* double complex */
  TTC_BF16 = 4, /* bfloat16 */
  TTC_F16 = 5,  /* IEEE binary16 */
  TTC_E4M3 = 6, /* OCP fp8 e4m3fn (output only) */
  TTC_E5M2 = 7  /* OCP fp8 e5m2 (output only) */
} ttc_dtype_t;

/* Status codes.  The reference has none: its wrapper printf()s and calls
 * exit(-1) (ttc/ttc_util.py:135-141); ttc_legacy_transpose keeps that
 * behaviour, every other entry point returns a status instead. */
typedef enum {
  TTC_SUCCESS = 0,
  TTC_ERROR_INVALID_VALUE = 1, /* bad perm/size/lda/ldb/pointer/flag */
  TTC_ERROR_UNSUPPORTED = 2,   /* type pair or shape the kernels do not cover */
  TTC_ERROR_HIP = 3,           /* a HIP runtime call failed */
  TTC_ERROR_ALLOC = 4,         /* host or device allocation failed */
  TTC_ERROR_NO_DEVICE = 5,     /* no usable gfx950 device */
  TTC_ERROR_WORKSPACE = 6      /* the work table exceeds TTC_BATCH_WORKSPACE_BYTES (ttc_batch_execute_noalloc) */
} ttc_status_t;

typedef struct ttc_plan_s* ttc_plan_t;
typedef struct ttc_batch_s* ttc_batch_t;
typedef void* ttc_stream_t; /* a hipStream_t; NULL = the null stream */

/* Kernel configuration: the search space of the autotuner.  It replaces the
 * reference's (blocking, loop order, threadsPerBlock) candidates
 * (ttc/GPUtransposeGenerator.py:258-285, 328-349; ttc/CUDAtranspose.py:6-111).
 * Every field set to 0 selects the built-in heuristic.
 * Indices of `tile` refer to the FUSED dimensions reported by
 * ttc_plan_get_info. */
typedef struct {
  int tile[TTC_MAX_DIM]; /* tile extent per fused A dimension (all 0 = heuristic) */
  int vec_a;             /* elements per lane vector on the A (read) side; 0 = widest legal */
  int vec_b;             /* elements per lane vector on the B (write) side; 0 = widest legal */
  int wg_per_cu;         /* resident persistent work-groups per CU; 0 = occupancy limit */
  int nt_store;          /* non-temporal B stores (beta_is_zero only): 0 = auto (off), 1 = on, 2 = off */
  int grid_order;        /* 0 = auto (4), 1 = strided tile walk, 2 = contiguous tile ranges,
                            3 = strided walk with the work-groups of one XCD on consecutive tiles,
                            4 = as 3 within one contiguous share of the tiles per XCD */
  int run_a;             /* heuristic target run length on A's side, elements (0 = default) */
  int run_b;             /* heuristic target run length on B's side, elements (0 = default) */
  int family;            /* kernel family: 0 = auto, 1 = box tile, 2 = row x column tile,
                            3 = row copy (perm[0] == 0 after fusion only) */
  int tile_order;        /* tile numbering (row x column tiles): 0 = auto (B order), 1 = A order, 2 = B order,
                            3 / 4 = B order in groups of 8 / 16 tiles along A's leading tiled dimension */
  int lds_buffers;       /* LDS tile buffers (row x column tiles): 0 = auto (2), 1 = one (two barriers per
                            tile, more resident work-groups), 2 = two (one barrier per tile) */
} ttc_config_t;

/* What a plan resolved to (fused geometry, tile, launch shape). */
typedef struct {
  int rank;                      /* after index fusion */
  int perm[TTC_MAX_DIM];
  int size[TTC_MAX_DIM];
  int64_t stride_a[TTC_MAX_DIM]; /* element stride of fused A dim k in A */
  int64_t stride_b[TTC_MAX_DIM]; /* element stride of fused A dim k in B */
  int tile[TTC_MAX_DIM];
  int vec_a, vec_b;
  int threads;                   /* work-group size (256) */
  int slots_a, slots_b;          /* vectors per thread per tile, read / write side */
  int64_t tiles;                 /* number of tiles */
  int tile_volume;               /* elements per full tile */
  int lds_bytes;                 /* dynamic LDS per work-group */
  int lds_pad;                   /* LDS padding (elements) on B's stride-1 dimension */
  int nt_store;
  int grid_order;
  int wg_per_cu;
  int64_t elements;              /* prod(size) */
  int64_t bytes_algorithmic;     /* N*sA + N*sB*(beta_is_zero ? 1 : 2) */
  int type_a, type_b;
  int beta_is_zero;
  int run_a, run_b;              /* contiguous element runs of a full tile in A and B */
  int family;                    /* 1 = box tile kernel, 2 = row x column tile kernel, 3 = row copy */
  int rows, cols;                /* row x column tile: row and column counts (0 for box tiles) */
  int tile_order;                /* 1 = A order, 2 = B order, 3 / 4 grouped B order */
  int lds_buffers;               /* LDS tile buffers of the row x column kernel (1 for other families) */
  int lane_fixed;                /* row x column kernel: full tiles take the lane-fixed walk (1) or the general one */
} ttc_plan_info_t;

/* Library version string. */
const char* ttc_version(void);
/* Human-readable status. */
const char* ttc_status_string(ttc_status_t s);
/* Bytes per element of a type (0 for unknown). */
int ttc_dtype_size(ttc_dtype_t t);

/* Index fusion: merges index runs that are contiguous in both A and B.
 * Replaces ttc/ttc_util.py:214-276 (fuseIndices) — same rule: a run of
 * consecutive A indices that are also consecutive in perm is fused when all
 * but its last dimension are unpadded in A and in B.  Outputs have
 * *out_dim entries; lda/ldb outputs are the fused outer extents. */
ttc_status_t ttc_fuse_indices(int dim, const int* perm, const int* size, const int* lda, const int* ldb,
                              int* out_dim, int* out_perm, int* out_size, int* out_lda, int* out_ldb);

/* Plan creation (pure host work; needs no device).  Validates like
 * ttc/ttc.py:733-768 (ld >= size, perm is a permutation), fuses indices,
 * and picks tile / vector widths.  cfg may be NULL. */
ttc_status_t ttc_plan_create(ttc_plan_t* plan, ttc_dtype_t type_a, ttc_dtype_t type_b, int dim, const int* perm,
                             const int* size, const int* lda, const int* ldb, int beta_is_zero,
                             const ttc_config_t* cfg);
ttc_status_t ttc_plan_get_info(ttc_plan_t plan, ttc_plan_info_t* info);
ttc_status_t ttc_plan_destroy(ttc_plan_t plan);

/* Asynchronous execution on `stream`.  Replaces the kernel launch of the
 * reference host wrapper (ttc/CUDAtranspose.py:418-421) without its
 * per-call metadata malloc/copy/free and device synchronisation.  Safe to
 * call concurrently from several threads on one plan. */
ttc_status_t ttc_execute(ttc_plan_t plan, const void* A, void* B, double alpha, double beta, ttc_stream_t stream);

/* fp8 cast-transpose (no counterpart in the reference).  type_b of the plan
 * must be TTC_E4M3 or TTC_E5M2, type_a TTC_S, TTC_BF16 or TTC_F16, and the
 * plan made with beta_is_zero = 1.
 *   B[pi(i)] = sat_rne_fp8(fl32(fl32(alpha*A[i]) * scale)),
 *   amax     = max |fl32(alpha*A[i])|.
 * mode TTC_SCALE_DELAYED: one pass; scale = *d_scale_in if d_scale_in is
 *   non-NULL, else the host value `scale`; amax of this tensor is written to
 *   *d_amax_out (if non-NULL).
 * mode TTC_SCALE_CURRENT: two passes; pass 1 reduces amax of A into
 *   *d_amax_out (required), pass 2 uses scale = fp8_max / amax (1 if amax is
 *   0 or not finite).
 * *d_scale_inv_out (if non-NULL) receives 1/scale. */
typedef enum { TTC_SCALE_DELAYED = 0, TTC_SCALE_CURRENT = 1 } ttc_scale_mode_t;
ttc_status_t ttc_execute_fp8(ttc_plan_t plan, const void* A, void* B, double alpha, const float* d_scale_in,
                             float scale, float* d_amax_out, float* d_scale_inv_out, ttc_scale_mode_t mode,
                             ttc_stream_t stream);

/* Synchronous drop-in with the reference wrapper's semantics
 * (ttc/CUDAtranspose.py:395-429): runs on the null stream, waits for the
 * device, and on any error prints "Kernel ERROR in <name>: <reason>" and calls
 * exit(-1) (ttc/ttc_util.py:135-141).  Plans are cached per argument set.
 * The per-transposition symbols <p>CuTranspose_<perm>_<size>[_bz] that the
 * generator emits (ttc/CUDAtranspose.py:438-451, 473-527) forward here.
 * cfg and name may be NULL. */
void ttc_legacy_transpose(ttc_dtype_t type_a, ttc_dtype_t type_b, const void* A, void* B, double alpha,
                          double beta, int beta_is_zero, int dim, const int* perm, const int* size, const int* lda,
                          const int* ldb, const ttc_config_t* cfg, const char* name);

/* Batched grouped transpositions (no counterpart in the reference, which
 * runs one transposition per host call).  All ops run in ONE persistent
 * kernel driven by a device-side work table built at creation. */
typedef struct {
  ttc_dtype_t type_a, type_b; /* real or complex, A and B of the same class */
  int dim;
  int perm[TTC_MAX_INPUT_DIM];
  int size[TTC_MAX_INPUT_DIM];
  int lda[TTC_MAX_INPUT_DIM]; /* all 0 = dense */
  int ldb[TTC_MAX_INPUT_DIM]; /* all 0 = dense */
  const void* A;
  void* B;
  double alpha, beta;
  int beta_is_zero;
  ttc_config_t cfg;           /* per-op tile choice; zeros = heuristic */
} ttc_batch_op_t;

/* Creates the work table on the current device (allocates device memory).
 * wg_per_cu: work-group slices per CU in each group's grid (residency is
 * set by occupancy; more slices than resident work-groups balance uneven
 * ops); 0 = 96. */
ttc_status_t ttc_batch_create(ttc_batch_t* batch, int nops, const ttc_batch_op_t* ops, int wg_per_cu);
/* One grouped launch of every op. */
ttc_status_t ttc_batch_execute(ttc_batch_t batch, ttc_stream_t stream);
/* The same work through a hipGraph captured on first use on `stream` (which
 * must not be the null stream): per_op_loop = 0 captures the grouped launch,
 * 1 captures the per-op launch loop. */
ttc_status_t ttc_batch_graph_launch(ttc_batch_t batch, ttc_stream_t stream, int per_op_loop);
/* Per-op loop with the same kernels (one launch per op), for comparison. */
ttc_status_t ttc_batch_execute_loop(ttc_batch_t batch, ttc_stream_t stream);
/* Number of tiles / work-groups / algorithmic bytes of the grouped launch. */
ttc_status_t ttc_batch_get_info(ttc_batch_t batch, int64_t* tiles, int* workgroups, int64_t* bytes);
ttc_status_t ttc_batch_destroy(ttc_batch_t batch);

/* One grouped launch of `ops` that allocates no device memory.  No
 * counterpart in the reference: it is for callers that cannot call
 * hipMalloc (code under stream capture, tight serving loops, harnesses that
 * forbid allocation).  It does what ttc_batch_create + ttc_batch_execute +
 * ttc_batch_destroy do, with the same kernels and results, except that the
 * work table goes into a fixed device workspace the library owns, one per
 * GPU, of TTC_BATCH_WORKSPACE_BYTES, held until the process ends.  The table
 * is built on the host (host heap only) and copied with hipMemcpyAsync on
 * `stream`, ordered after that stream's earlier work; the call returns once
 * the kernels are enqueued.  A table larger than the workspace (about 24
 * bytes per tile plus about 1 KiB per op) returns TTC_ERROR_WORKSPACE.
 * Every call on one GPU shares that GPU's workspace: calls on one stream
 * are ordered, but calls that overlap on one GPU (different streams or host
 * threads, without ordering between them) race. */
#define TTC_BATCH_WORKSPACE_BYTES (16 << 20)
ttc_status_t ttc_batch_execute_noalloc(int nops, const ttc_batch_op_t* ops, int wg_per_cu, ttc_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TTC_HIP_H */
