/* C ABI of the gfx950 SPAI-GMRES-IR library (libbspai_hip.so).
 *
 * Plain C types only: pointers, sizes, int status codes; no C++ or torch types
 * cross this boundary and no exception escapes it.  Calls that take host
 * arrays, and every solver call (they report statuses), are synchronous when
 * they return, as the reference's CUDA objects are
 * (include/bspai/cuda_gmres.hpp:18-20); bspai_hip_spmv / _spmv_f32 / _sddmm on
 * device arrays are stream-ordered on the context's stream.  A context is not
 * thread-safe.
 *
 * Vectors and matrices may live on the host (where = BSPAI_HIP_HOST: copied
 * in/out by the library) or on the device (BSPAI_HIP_DEVICE: used in place;
 * e.g. torch tensor data pointers).  Indices are int32 on the device side.
 *
 * Reference interfaces replaced (reference tree, file:line):
 *   bspai_hip_matrix_create / bspai_hip_spmv    CudaCsrSpmv      include/bspai/cuda_spmv.hpp:7-19
 *   bspai_hip_csr_from_entries                  CsrMatrix::from_entries include/bspai/cpu.hpp:19
 *   bspai_hip_mm_parse (+ _info/_entries)       read_matrix_market include/bspai/cpu.hpp:47
 *   bspai_hip_mm_parse_gpu (+ _gpu_info/_entries/_copy)   the same, parsed on the GPU
 *                                               (the parse; canonicalization is
 *                                               bspai_hip_csr_from_entries)
 *   bspai_hip_spai_build (+ _info/_columns/...) cuda_construct_spai include/bspai/cuda_spai.hpp:47
 *                                               construct_spai    include/bspai/spai.hpp:69
 *   bspai_hip_score                             cuda_evaluate_spai_score include/bspai/cuda_spai.hpp:48-49
 *   bspai_hip_gmres                             cuda_left_preconditioned_gmres include/bspai/cuda_gmres.hpp:40-45
 *                                               left_preconditioned_gmres include/bspai/gmres.hpp:55-57
 *   bspai_hip_ir_solve                          cuda_refine_with_preconditioner include/bspai/cuda_ir.hpp:19-24
 *   bspai_hip_spai_gmres_ir                     cuda_gpu_spai_gmres_ir include/bspai/cuda_ir.hpp:26-28
 *                                               spai_gmres_ir     include/bspai/ir.hpp:65-66
 *   bspai_hip_adjoint_solve, bspai_hip_sddmm    no reference counterpart (differentiable solve)
 *   bspai_hip_ctx_set_timing / _kernel_times    telemetry: the reference's GPU results carry
 *                                               CudaSpaiTelemetry include/bspai/cuda_spai.hpp:9-37
 *                                               and CudaIrResult counters include/bspai/cuda_ir.hpp:7-16
 */
#ifndef BSPAI_HIP_H
#define BSPAI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSPAI_HIP_ABI_VERSION 1

/* return codes */
#define BSPAI_HIP_OK 0
#define BSPAI_HIP_EINVAL 1  /* invalid argument (message in bspai_hip_last_error) */
#define BSPAI_HIP_EDEVICE 2 /* HIP runtime / device failure */
#define BSPAI_HIP_ENOMEM 3  /* device allocation failed */
#define BSPAI_HIP_EFORMAT 4 /* malformed Matrix Market text (message: the reference reader's) */

/* memory kinds */
#define BSPAI_HIP_HOST 0
#define BSPAI_HIP_DEVICE 1

/* precision of the SPAI build and of the inner GMRES */
#define BSPAI_HIP_FP64 0  /* plain: FP64 SPAI, GMRES and IR */
#define BSPAI_HIP_MIXED 1 /* FP32 SPAI + FP32 GMRES inner, FP64 residual/update/Givens */

/* status values mirror the reference enums:
 *   SpaiStatus  include/bspai/spai.hpp:10-21   (success=0 ... iteration_limit=9)
 *   GmresStatus include/bspai/gmres.hpp:10-20  (success=0 ... numerical_failure=8)
 *   IrStatus    include/bspai/ir.hpp:9-17      (success=0 ... numerical_failure=6) */

#define BSPAI_HIP_MAX_OUTER 24

typedef struct bspai_hip_ctx bspai_hip_ctx;
typedef struct bspai_hip_matrix bspai_hip_matrix;
typedef struct bspai_hip_spai bspai_hip_spai;

typedef struct {
    double epsilon; /* accept a column when ||B(I,J) m - e_k|| <= epsilon */
    int64_t alpha;  /* growth budget per column */
    int64_t beta;   /* max indices added per growth */
    int precision;  /* BSPAI_HIP_FP64 or BSPAI_HIP_MIXED */
} bspai_hip_spai_options;

typedef struct {
    int status;              /* SpaiStatus */
    int64_t failing_column;  /* first failing column, -1 if none */
    int64_t n, p_nnz, pattern_nnz;
    int64_t fits, growths, clamps;
    int passes;              /* capacity buckets used */
    double build_ms;         /* device time of the construction */
    double qr_flops;         /* sum over accepted fits of 2mq^2 - 2q^3/3 + 4mq + q^2 */
    char message[256];
} bspai_hip_spai_info;

typedef struct {
    double tolerance;        /* tau on ||P(b-Ax)||/||Pb|| */
    int64_t max_iterations;  /* 0 = n */
} bspai_hip_gmres_options;

typedef struct {
    int status, converged;
    int64_t iterations;
    double initial_original, initial_relative, final_original, final_relative;
} bspai_hip_gmres_info;

typedef struct {
    double spai_epsilon;
    int64_t spai_alpha;      /* 0 = ceil(n/8), as the solver CLI */
    int64_t spai_beta;
    double gmres_tolerance;  /* inner tau; BSPAI_HIP_MIXED: the floor of an
                                adaptive tau, max(gmres_tolerance, 0.3 x the
                                contraction the outer stop still needs) */
    int64_t gmres_max_iterations; /* 0 = n */
    int64_t max_outer_steps; /* <= 20 */
    int precision;           /* BSPAI_HIP_FP64 or BSPAI_HIP_MIXED */
    double relres_target;    /* > 0: also require ||b-Ax||_2 <= target*||b||_2 */
} bspai_hip_ir_options;

typedef struct {
    int status, converged, construction_performed, spai_status;
    int64_t outer_steps, inner_attempts;
    int inner_status[BSPAI_HIP_MAX_OUTER];
    int inner_iterations[BSPAI_HIP_MAX_OUTER];
    int history_len;
    double history[BSPAI_HIP_MAX_OUTER][4]; /* residual_inf, residual_2, nbe, cbe */
    double residual_inf, residual_2, normwise_backward_error, componentwise_backward_error;
    double spai_ms, solve_ms;
    char message[256];
} bspai_hip_ir_info;

int bspai_hip_abi_version(void);

/* context: device ordinal; stream = hipStream_t to run on, or NULL for a
 * library-owned stream (created at default priority). */
int bspai_hip_ctx_create(int device, void *stream, bspai_hip_ctx **out);
void bspai_hip_ctx_destroy(bspai_hip_ctx *ctx);
const char *bspai_hip_last_error(const bspai_hip_ctx *ctx);
int bspai_hip_synchronize(bspai_hip_ctx *ctx);

/* Kernel timing telemetry.  While enabled, each launch of the hot kernels
 * (SPAI columns, the two SpMVs and the Gram-Schmidt passes of every Arnoldi
 * step) is bracketed by two HIP events on the context's stream; enable = 2
 * times the SPAI construction kernels only (two events per capacity pass,
 * negligible overhead), enable = 1 every hot kernel (a few percent of a
 * GMRES step), 0 stops.
 * bspai_hip_ctx_kernel_times synchronizes and returns the totals per kernel
 * label since timing was last enabled: it writes at most `capacity` entries
 * and sets *count to the number of labels.  Enabling or disabling clears
 * the totals. */
typedef struct {
    char name[96];       /* kernel name as rocprofv3 lists it, plus the operand role */
    int64_t launches;
    double total_ms;     /* sum of the event-measured durations */
    double bytes;        /* algorithmic HBM bytes, summed over the launches */
    double flops;        /* algorithmic flops, summed over the launches */
} bspai_hip_kernel_time;
int bspai_hip_ctx_set_timing(bspai_hip_ctx *ctx, int enable);
/* Arnoldi-step graphs (hipGraph) the context's GMRES solves have captured and
 * replayed so far.  Each replay runs several Arnoldi steps between two host
 * check points; BSPAI_GMRES_GRAPH=0 in the environment, or timing enabled,
 * launches the same kernels one by one on the stream instead. */
int bspai_hip_ctx_graph_stats(const bspai_hip_ctx *ctx, int64_t *captured, int64_t *replayed);
int bspai_hip_ctx_kernel_times(bspai_hip_ctx *ctx, bspai_hip_kernel_time *out, int capacity,
                               int *count);

/* n x n CSR (canonical: sorted columns, no duplicates).  nsys > 1 declares a
 * block-diagonal batch of independent systems: system s owns rows
 * [sys_offsets[s], sys_offsets[s+1]) (host array, nsys+1 entries). */
int bspai_hip_matrix_create(bspai_hip_ctx *ctx, int64_t n, int64_t nnz, const int32_t *ptr,
                            const int32_t *col, const double *val, int where, int nsys,
                            const int64_t *sys_offsets, bspai_hip_matrix **out);
void bspai_hip_matrix_destroy(bspai_hip_matrix *m);
int bspai_hip_matrix_info(const bspai_hip_matrix *m, int64_t *n, int64_t *nnz, int *nsys);
int bspai_hip_matrix_set_values(bspai_hip_ctx *ctx, bspai_hip_matrix *m, const double *val,
                                int where);
int bspai_hip_matrix_download(bspai_hip_ctx *ctx, const bspai_hip_matrix *m, int32_t *ptr,
                              int32_t *col, double *val);
int bspai_hip_matrix_transpose(bspai_hip_ctx *ctx, const bspai_hip_matrix *m,
                               bspai_hip_matrix **out);

/* Canonical rows x cols CSR from `count` coordinate entries (row, col, value),
 * on the device.  Replaces CsrMatrix::from_entries (reference
 * include/bspai/cpu.hpp:19, src/cpu.cpp:48-76) with the same result: entries
 * sorted by (row, column), duplicates summed in input order, exact zeros
 * dropped.  Errors (BSPAI_HIP_EINVAL, with the reference's messages): an index
 * outside the matrix or a non-finite value ("Entry index outside matrix
 * dimensions" / "Matrix entries must be finite", the first offending entry in
 * input order deciding, its indices checked before its value), then a
 * duplicate sum that becomes non-finite ("Duplicate sum is nonfinite").
 * row, col, val: `count` entries in `where` memory.  Outputs in `out_where`
 * memory: ptr (rows + 1 entries), col_out and val_out (room for `count`; the
 * first *nnz are written; after an error their contents are unspecified).
 * rows, cols and count below 2^31 - 1.  Device scratch (about 60 bytes per
 * entry) is kept by the context for later calls. */
int bspai_hip_csr_from_entries(bspai_hip_ctx *ctx, int64_t rows, int64_t cols, int64_t count,
                               const int64_t *row, const int64_t *col, const double *val,
                               int where, int32_t *ptr, int32_t *col_out, double *val_out,
                               int out_where, int64_t *nnz);

/* Allocation-free variant of bspai_hip_csr_from_entries.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result, messages and error order as
 * bspai_hip_csr_from_entries; every array (row, col, val in; ptr, col_out,
 * val_out out) must be device memory.  It allocates nothing: its scratch is a
 * static device array of 48 MiB per GPU, held until the process ends, whose
 * address bspai_hip_ctx_create looks up once per device.  Limits:
 * count <= 524288 (2^19) entries and rows < 4194304 (2^22); larger inputs
 * return BSPAI_HIP_EINVAL.  Every call on one GPU shares that array: calls
 * that overlap on one GPU (several threads or streams) race.  Synchronous. */
int bspai_hip_csr_from_entries_noalloc(bspai_hip_ctx *ctx, int64_t rows, int64_t cols,
                                       int64_t count, const int64_t *row, const int64_t *col,
                                       const double *val, int32_t *ptr, int32_t *col_out,
                                       double *val_out, int64_t *nnz);

/* Matrix Market coordinate reader (host code, several threads; no GPU use).
 * Parses `len` bytes of text with the reference reader's rules and messages
 * (read_matrix_market, src/cpu.cpp:105-161 of the reference): banner
 * "%%MatrixMarket matrix coordinate real|integer general|symmetric", '%'
 * comment and blank lines skipped, one dimension line, exactly `count` entry
 * lines of "row col value" (one-based), nothing after them.  On success *out
 * holds the entries in file order, zero-based, a symmetric file's off-diagonal
 * entries followed by their mirror; they are not canonicalized (duplicates
 * and zeros kept): pass them to bspai_hip_csr_from_entries.  A malformed file
 * returns BSPAI_HIP_EFORMAT with the reference's message (a std::runtime_error
 * there) in `message` (message_len bytes, terminated).  threads: 0 = up to 16,
 * else that many (the body is split at line breaks). */
typedef struct bspai_hip_mm bspai_hip_mm;
int bspai_hip_mm_parse(const char *text, int64_t len, int threads, bspai_hip_mm **out,
                       char *message, int message_len);
int bspai_hip_mm_info(const bspai_hip_mm *mm, int64_t *rows, int64_t *cols, int64_t *count,
                      int *integer_field, int *symmetric);
/* row, col, val: room for `count` entries each (host arrays). */
int bspai_hip_mm_entries(const bspai_hip_mm *mm, int64_t *row, int64_t *col, double *val);
void bspai_hip_mm_destroy(bspai_hip_mm *mm);

/* The same parse on the GPU (replaces read_matrix_market, src/cpu.cpp:105-161,
 * like bspai_hip_mm_parse; the reference reader is host code): the banner and
 * the dimension line are read on the host, the body is copied to the device
 * once and its lines are found, parsed and put in file order there, with the
 * same rules, the same values bit for bit and the same messages (BSPAI_HIP_EFORMAT).
 * The entries stay on the device until bspai_hip_mm_gpu_destroy.  text: host
 * bytes; bodies of up to 2^31 - 3 lines. */
typedef struct bspai_hip_mm_gpu bspai_hip_mm_gpu;
int bspai_hip_mm_parse_gpu(bspai_hip_ctx *ctx, const char *text, int64_t len,
                           bspai_hip_mm_gpu **out, char *message, int message_len);
int bspai_hip_mm_gpu_info(const bspai_hip_mm_gpu *mm, int64_t *rows, int64_t *cols,
                          int64_t *count, int *integer_field, int *symmetric);
/* Device pointers to the `count` entries (zero-based int64 indices, FP64
 * values), valid until bspai_hip_mm_gpu_destroy: pass them to
 * bspai_hip_csr_from_entries with BSPAI_HIP_DEVICE. */
int bspai_hip_mm_gpu_entries(const bspai_hip_mm_gpu *mm, const int64_t **row, const int64_t **col,
                             const double **val);
/* The entries copied into caller arrays in `where` memory. */
int bspai_hip_mm_gpu_copy(bspai_hip_ctx *ctx, const bspai_hip_mm_gpu *mm, int64_t *row,
                          int64_t *col, double *val, int where);
void bspai_hip_mm_gpu_destroy(bspai_hip_mm_gpu *mm);

/* y = A x in FP64.  lanes: 0 = the matrix's default; 1 = streaming kernel
 * (the block's entries staged in LDS); -1 = one thread per row; 2..64 (a
 * power of two) = that many lanes per row; BSPAI_HIP_SPMV_SLICED = one
 * wavefront per 64-row slice of a sliced copy (entries column-major per
 * slice, int32 columns), BSPAI_HIP_SPMV_SLICED16 = the same on int16 column
 * offsets (both EINVAL when a row has more than 255 entries or the slices
 * would hold more than 3x the entries; SLICED16 also when an offset
 * col - row does not fit 16 bits).  1, -1 and the sliced
 * kernels add each row in CSR order, bit-identical to the reference's
 * CsrMatrix::multiply. */
#define BSPAI_HIP_SPMV_SLICED 256
#define BSPAI_HIP_SPMV_SLICED16 257
int bspai_hip_spmv(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const double *x, double *y,
                   int where, int lanes);
/* y = A x with FP32 values and vectors (the mixed path's inner product). */
int bspai_hip_spmv_f32(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const float *x, float *y,
                       int where, int lanes);

/* SPAI construction (device resident), P = N^T D. */
int bspai_hip_spai_build(bspai_hip_ctx *ctx, const bspai_hip_matrix *a,
                         const bspai_hip_spai_options *opt, bspai_hip_spai **out);
void bspai_hip_spai_destroy(bspai_hip_spai *p);
int bspai_hip_spai_get_info(const bspai_hip_spai *p, bspai_hip_spai_info *info);
/* canonical P (int32 structure, values widened to FP64), host arrays:
 * ptr[n+1], col[p_nnz], val[p_nnz] */
int bspai_hip_spai_preconditioner(bspai_hip_ctx *ctx, const bspai_hip_spai *p, int32_t *ptr,
                                  int32_t *col, double *val);
/* per-column diagnostics, host arrays of n */
int bspai_hip_spai_columns(bspai_hip_ctx *ctx, const bspai_hip_spai *p, int32_t *status,
                           int32_t *fits, int32_t *growths, int32_t *clamps, double *residual,
                           double *pivot_ratio);
/* final patterns J_k and coefficients (exact zeros kept): ptr[n+1] (int64),
 * idx[pattern_nnz], coef[pattern_nnz] */
int bspai_hip_spai_patterns(bspai_hip_ctx *ctx, const bspai_hip_spai *p, int64_t *ptr,
                            int32_t *idx, double *coef);
int bspai_hip_spai_scaling(bspai_hip_ctx *ctx, const bspai_hip_spai *p, double *d);
/* restricted score rho of one candidate on the device (FP64). */
int bspai_hip_score(bspai_hip_ctx *ctx, double residual_squared, double dot,
                    double denominator_squared, int *status, int *clamped, double *score);

/* One left-preconditioned GMRES solve (FP64) per system; info[nsys]. */
int bspai_hip_gmres(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const bspai_hip_matrix *p,
                    const double *rhs, const double *x0, const bspai_hip_gmres_options *opt,
                    double *x, int where, bspai_hip_gmres_info *info);

/* Iterative refinement with a given preconditioner P (any matrix handle, e.g.
 * bspai_hip_spai_matrix of a build); info[nsys]. */
int bspai_hip_ir_solve(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const bspai_hip_matrix *p,
                       const double *b, const bspai_hip_ir_options *opt, double *x, int where,
                       bspai_hip_ir_info *info);
/* SPAI construction + refinement (the solver CLI's call); info[nsys]. */
int bspai_hip_spai_gmres_ir(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const double *b,
                            const bspai_hip_ir_options *opt, double *x, int where,
                            bspai_hip_ir_info *info);

/* Adjoint solve A^T lambda = g with the transposed preconditioner (right
 * preconditioning by P^T) under the same IR/precision options; at = A^T and
 * pt = P^T as matrices; info[nsys]. */
int bspai_hip_adjoint_solve(bspai_hip_ctx *ctx, const bspai_hip_matrix *at,
                            const bspai_hip_matrix *pt, const double *g,
                            const bspai_hip_ir_options *opt, double *lambda, int where,
                            bspai_hip_ir_info *info);
/* P (or P^T) of a build as a matrix handle (FP64 values). */
int bspai_hip_spai_matrix(bspai_hip_ctx *ctx, const bspai_hip_spai *p, int transpose,
                          bspai_hip_matrix **out);
/* out[q] = alpha * u[row(q)] * v[col(q)] on A's stored pattern (SDDMM). */
int bspai_hip_sddmm(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, double alpha, const double *u,
                    const double *v, double *out, int where);

#ifdef __cplusplus
}
#endif
#endif
