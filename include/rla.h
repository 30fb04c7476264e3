/*
 * rla.h — C ABI of librla_mi355x.so, the gfx950 (MI355X) kernels behind the
 * RandLinearAlgebra.jl compressor / approximator / solver surface.
 *
 * Conventions (all entry points):
 *   - Return an int status (RLA_OK on success).  On failure a thread-local
 *     message is available from rla_last_error().  Status codes map to the
 *     reference's exception classes: RLA_EARG -> ArgumentError,
 *     RLA_EDIM -> DimensionMismatch, RLA_EDOMAIN -> DomainError.
 *   - Matrices are passed as (pointer, row stride, column stride) in elements,
 *     so a Julia column-major matrix is (p, 1, ld) and its transpose/adjoint
 *     view is (p, ld, 1).  Pointers are device pointers unless the name says
 *     "host".
 *   - Index arrays are 0-based int32 on the device.  Julia's 1-based Int64
 *     arrays are converted at the boundary with rla_index_from_julia, which
 *     range-checks them.
 *   - Work is enqueued on the context's stream; nothing blocks unless the
 *     entry point says so.  Reads of A are borrowed, never retained.
 *   - One context per host thread: a context owns a stream and a workspace,
 *     and two calls that share a context must not overlap.
 */
#ifndef RLA_MI355X_H
#define RLA_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RLA_OK = 0,
  RLA_EARG = 1,     /* ArgumentError: bad parameter / unsupported combination */
  RLA_EDIM = 2,     /* DimensionMismatch (messages mirror src/helpers/mul_dim_checks.jl:41-159) */
  RLA_EDOMAIN = 3,  /* DomainError */
  RLA_ENOMEM = 4,   /* device allocation failed */
  RLA_EADMIT = 5,   /* request over the device's HBM budget (rla_hbm_*) */
  RLA_EHIP = 6,     /* HIP runtime error */
  RLA_ENOTSUP = 7   /* size outside what the kernels support */
};

typedef struct rla_ctx rla_ctx;

/* ---- context --------------------------------------------------------- */
const char* rla_version(void);
const char* rla_last_error(void);
/* Work goes to hip_stream; NULL is the device's null stream (what torch's
 * default stream is). */
int rla_ctx_create(int device, void* hip_stream, rla_ctx** out);
int rla_ctx_destroy(rla_ctx* ctx);
/* Later calls go to hip_stream.  Work the context enqueued on its previous
 * stream is ordered before them (an event on the old stream that the new one
 * waits for), since it may still use the context's workspace. */
int rla_ctx_set_stream(rla_ctx* ctx, void* hip_stream);

/* ---- HBM admission ----------------------------------------------------
 * One account per device, shared by every context of the process.  No
 * counterpart in the reference (its Julia arrays are host memory); it is the
 * admission-by-HBM-budget of a serving loop, moved below the C boundary so
 * that a caller of the library gets it without the Python host layer.
 *   budget 0 (the default) means unlimited.
 *   rla_hbm_admit charges `bytes` for a request (its resident operands and
 *   working memory) and fails with RLA_EADMIT, charging nothing, when the
 *   account would exceed the budget; rla_hbm_release returns them.
 *   The library's own device allocations (context workspaces, batched
 *   least-squares plans) are charged to the same account: an entry point
 *   whose workspace would have to grow past the budget fails with
 *   RLA_EADMIT before it allocates or launches anything.
 * These calls touch no device and are safe from any thread. */
int rla_hbm_set_budget(int device, size_t bytes);
int rla_hbm_admit(int device, size_t bytes);
int rla_hbm_release(int device, size_t bytes);
int rla_hbm_query(int device, size_t* used, size_t* budget);
int rla_ctx_sync(rla_ctx* ctx);
/* Bytes of device workspace currently held by the context. */
int rla_ctx_workspace_bytes(rla_ctx* ctx, size_t* out);

/* ---- counter-based generator ------------------------------------------
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Host evaluation, for
 * known-answer tests.  Every index, sign and Gaussian draw of the recipe
 * generators below is a pure function of (seed, epoch, item), identical on
 * host and device.  Replaces the Random/StatsBase draws at
 * src/Compressors/sparse_sign.jl:167,252,317, count_sketch.jl:124-125,160-161,
 * srht.jl:150-151,207-208, gaussian.jl:107,122,138. */
void rla_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Julia 1-based Int64 index vector -> 0-based int32 (device), each entry
 * checked to lie in 1:maxval (RLA_EARG otherwise; blocks to read the check). */
int rla_index_from_julia(rla_ctx* ctx, const int64_t* idx1, int64_t n, int64_t maxval,
                         int32_t* idx0);

/* SparseSign recipe: for each of n_groups groups, nnz distinct sorted rows
 * from 0:s-1 (StatsBase sample!(..., replace=false, ordered=true), as in
 * sparse_idx_update!, src/Compressors/sparse_sign.jl:155-172) and values
 * +-1/sqrt(nnz) in the element type (sparse_sign.jl:248-252).
 * idx and val hold n_groups*nnz entries, group-major (CSC column order). */
int rla_sparse_sign_gen_f32(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                            int64_t s, int64_t nnz, int32_t* idx, float* val);
int rla_sparse_sign_gen_f64(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                            int64_t s, int64_t nnz, int32_t* idx, double* val);
int rla_host_sparse_sign_gen_f64(uint64_t seed, uint64_t epoch, int64_t n_groups, int64_t s,
                                 int64_t nnz, int32_t* idx_host, double* val_host);

/* CountSketch recipe: one row uniform in 0:s-1 and one sign +-1 per group
 * (src/Compressors/count_sketch.jl:124-127,160-161). */
int rla_count_sketch_gen_f32(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                             int64_t s, int32_t* idx, float* val);
int rla_count_sketch_gen_f64(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_groups,
                             int64_t s, int32_t* idx, double* val);
int rla_host_count_sketch_gen_f64(uint64_t seed, uint64_t epoch, int64_t n_groups, int64_t s,
                                  int32_t* idx_host, double* val_host);

/* SRHT/FJLT recipe: sign bits for n_pad entries in Julia BitVector chunk
 * layout (bit i of chunk i>>6, LSB first; srht.jl:150 bitrand) and, when
 * op != NULL, s distinct sampled rows of 0:n_pad-1 in random order
 * (srht.jl:151, sample(1:n_pad, s, replace=false)). */
int rla_srht_gen(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n_pad, int64_t s,
                 int32_t* op, uint64_t* sign_chunks);
int rla_host_srht_gen(uint64_t seed, uint64_t epoch, int64_t n_pad, int64_t s,
                      int32_t* op_host, uint64_t* sign_chunks_host);

/* Gaussian recipe: out[i + j*ld] = scale * N(0,1) draw (gaussian.jl:107,122). */
int rla_gaussian_gen_f32(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t rows,
                         int64_t cols, double scale, float* out, int64_t ld);
int rla_gaussian_gen_f64(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t rows,
                         int64_t cols, double scale, double* out, int64_t ld);
int rla_host_gaussian_gen_f64(uint64_t seed, uint64_t epoch, int64_t rows, int64_t cols,
                              double scale, double* out_host, int64_t ld);

/* ---- sparse compressors (SparseSign, CountSketch, FJLT's K) ------------
 * M is an nrows x ncols CSC matrix: colptr[ncols+1], rowidx[nnz], val[nnz],
 * rows sorted within each column.  rla_sparse_transpose builds its CSR twin
 * (rowptr[nrows+1], colidx[nnz] ascending within each row, val[nnz]);
 * deterministic.  Replaces the SparseMatrixCSC the recipes hold
 * (sparse_sign.jl:219-295, count_sketch.jl:105-147). */
int rla_sparse_transpose_f32(rla_ctx* ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                             const int32_t* colptr, const int32_t* rowidx, const float* val,
                             int32_t* rowptr, int32_t* colidx, float* val_t);
int rla_sparse_transpose_f64(rla_ctx* ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                             const int32_t* colptr, const int32_t* rowidx, const double* val,
                             int32_t* rowptr, int32_t* colidx, double* val_t);

/* C = alpha * op(M) * X + beta * C, op(M) = M (trans_m = 0) or M' (trans_m = 1).
 * X is op(M)-cols x ncols, C is op(M)-rows x ncols, both strided.  Every
 * orientation of the reference's sparse mul! reduces to this one by taking
 * transposed views:  A*M = (M'*A')',  A*M' = (M*A')'.
 * Replaces mul! at sparse_sign.jl:339-361 and count_sketch.jl:166-213 (and
 * through them the SparseArrays sparse x dense kernels).  Accumulation per
 * output entry: c = beta*c (or 0 when beta == 0), then c += v*(x*alpha) over
 * the nonzeros in ascending index order — the SparseArrays loop order. */
int rla_sparse_apply_f32(rla_ctx* ctx, int trans_m, int64_t nrows, int64_t ncols,
                         const int32_t* colptr, const int32_t* rowidx, const float* val,
                         const int32_t* rowptr, const int32_t* colidx, const float* val_t,
                         int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs,
                         float* C, int64_t c_rs, int64_t c_cs, double alpha, double beta);
int rla_sparse_apply_f64(rla_ctx* ctx, int trans_m, int64_t nrows, int64_t ncols,
                         const int32_t* colptr, const int32_t* rowidx, const double* val,
                         const int32_t* rowptr, const int32_t* colidx, const double* val_t,
                         int64_t n_rhs, const double* X, int64_t x_rs, int64_t x_cs,
                         double* C, int64_t c_rs, int64_t c_cs, double alpha, double beta);

/* Row-interleaved copy of the gather structure of op(M) (its CSR for op = M,
 * its CSC for op = M'; `rows` = rows of op(M), ptr/idx/val those arrays):
 * entry q of row i at ell_idx/ell_val[q * rows + i], q < width, padded with
 * index 0 and value 0 past the row length.  rla_sparse_ell_width returns the
 * longest row (synchronises).  The layout is what the lanes of a wavefront,
 * one output row each, read as consecutive words; built once per recipe
 * (after complete_compressor / update_compressor!, sparse_sign.jl:229-336,
 * count_sketch.jl:113-162). */
int rla_sparse_ell_width(rla_ctx* ctx, int64_t rows, const int32_t* ptr, int32_t* width);
int rla_sparse_ell_build_f32(rla_ctx* ctx, int64_t rows, int32_t width, const int32_t* ptr,
                             const int32_t* idx, const float* val, int32_t* ell_idx,
                             float* ell_val);
int rla_sparse_ell_build_f64(rla_ctx* ctx, int64_t rows, int32_t width, const int32_t* ptr,
                             const int32_t* idx, const double* val, int32_t* ell_idx,
                             double* ell_val);

/* C = alpha * op(M) * X + beta * C with op(M) given by its row lengths (ptr,
 * rows + 1 entries) and the row-interleaved arrays above; X column-major
 * (x_rs == 1), in_rows x n_rhs.  Same arithmetic and order as
 * rla_sparse_apply, so results are bit-identical to it.  Replaces mul! at
 * sparse_sign.jl:339-361 and count_sketch.jl:166-213 for column-major X. */
int rla_sparse_apply_ell_f32(rla_ctx* ctx, int64_t out_rows, int64_t in_rows, const int32_t* ptr,
                             const int32_t* ell_idx, const float* ell_val, int64_t n_rhs,
                             const float* X, int64_t x_rs, int64_t x_cs, float* C, int64_t c_rs,
                             int64_t c_cs, double alpha, double beta);
int rla_sparse_apply_ell_f64(rla_ctx* ctx, int64_t out_rows, int64_t in_rows, const int32_t* ptr,
                             const int32_t* ell_idx, const double* ell_val, int64_t n_rhs,
                             const double* X, int64_t x_rs, int64_t x_cs, double* C,
                             int64_t c_rs, int64_t c_cs, double alpha, double beta);

/* Sign-packed row-interleaved copy, for operators whose nonzeros share one
 * magnitude (SparseSign +-1/sqrt(nnz), sparse_sign.jl:248-252; CountSketch
 * +-1, count_sketch.jl:124-127): entry q of row i is the 32-bit word
 * column | (value < 0) << 31 at ell_sidx[q * rows + i], 0 past the row length;
 * *magnitude receives the common |value| (host pointer; synchronises).
 * RLA_EDOMAIN if two nonzeros differ in magnitude. */
int rla_sparse_ell_sign_build_f32(rla_ctx* ctx, int64_t rows, int32_t width,
                                  const int32_t* ptr, const int32_t* idx, const float* val,
                                  uint32_t* ell_sidx, double* magnitude);
int rla_sparse_ell_sign_build_f64(rla_ctx* ctx, int64_t rows, int32_t width,
                                  const int32_t* ptr, const int32_t* idx, const double* val,
                                  uint32_t* ell_sidx, double* magnitude);

/* The same product as rla_sparse_apply_ell (and as mul! at
 * sparse_sign.jl:339-361, count_sketch.jl:166-213) from the sign-packed copy:
 * one workgroup per column of X stages the whole column in LDS, so
 * 1 <= in_rows <= 160 KiB / element size (else RLA_ENOTSUP).  Same arithmetic
 * and order: bit-identical to rla_sparse_apply. */
int rla_sparse_apply_ell_sign_f32(rla_ctx* ctx, int64_t out_rows, int64_t in_rows,
                                  const int32_t* ptr, const uint32_t* ell_sidx, double magnitude,
                                  int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs,
                                  float* C, int64_t c_rs, int64_t c_cs, double alpha,
                                  double beta);
int rla_sparse_apply_ell_sign_f64(rla_ctx* ctx, int64_t out_rows, int64_t in_rows,
                                  const int32_t* ptr, const uint32_t* ell_sidx, double magnitude,
                                  int64_t n_rhs, const double* X, int64_t x_rs, int64_t x_cs,
                                  double* C, int64_t c_rs, int64_t c_cs, double alpha,
                                  double beta);

/* Block-interleaved copy of op(M) for X taller than one LDS block: the rows of
 * X are cut into blocks of rb rows; entry s (in column order) of row i among
 * its columns in block b is at bidx/bval[rows * (w_0 + ... + w_{b-1}) + s * rows
 * + i], w_b being the most entries any row has in block b, and bidx = -1 past
 * a row's entries.  Every load of the gather below then reads consecutive
 * words.  No counterpart in the reference; it serves the same mul! as
 * rla_sparse_apply (fjlt.jl:257-319 through K, sparse_sign.jl:339-361).
 * rla_sparse_bell_widths writes w (ceil(in_rows / rb) device int32) and
 * *total_slots = sum of w (host pointer; synchronises); the copy holds
 * total_slots * rows entries.  RLA_EDOMAIN if a row's columns are outside
 * [0, in_rows) or not strictly ascending; RLA_ENOTSUP beyond 2^31 entries. */
int rla_sparse_bell_widths(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                           const int32_t* ptr, const int32_t* idx, int32_t* widths,
                           int64_t* total_slots);
int rla_sparse_bell_build_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                              const int32_t* ptr, const int32_t* idx, const float* val,
                              const int32_t* widths, int32_t* bidx, float* bval);
int rla_sparse_bell_build_f64(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                              const int32_t* ptr, const int32_t* idx, const double* val,
                              const int32_t* widths, int32_t* bidx, double* bval);

/* C = alpha * op(M) * X + beta * C from the block-interleaved copy; X
 * column-major (x_rs == 1), in_rows x n_rhs; rb * element size <= 64 KiB
 * (else RLA_ENOTSUP).  Same arithmetic and order as rla_sparse_apply
 * (sparse_sign.jl:339-361), so results are bit-identical to it. */
int rla_sparse_apply_bell_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                              const int32_t* widths, const int32_t* bidx, const float* bval,
                              int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs, float* C,
                              int64_t c_rs, int64_t c_cs, double alpha, double beta);
int rla_sparse_apply_bell_f64(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                              const int32_t* widths, const int32_t* bidx, const double* bval,
                              int64_t n_rhs, const double* X, int64_t x_rs, int64_t x_cs,
                              double* C, int64_t c_rs, int64_t c_cs, double alpha, double beta);

/* The same product for tall X by runs of blocks: a workgroup takes up to 8
 * columns of X (as many rb-row blocks as fit 128 KiB of LDS; each entry of M
 * it reads serves all of them) and, for at most 1024 output rows, a run of
 * consecutive blocks, writing each row's FP64 sum over its run; the runs are
 * added in ascending order, so the result is within rounding of
 * rla_sparse_apply (not identical: sums are formed per run).  Smaller blocks
 * (rb) let more columns share the entries read.  The context workspace holds
 * the runs' sums: at most rla_sparse_apply_bell_split_bytes bytes.
 * Replaces the same mul! as rla_sparse_apply_bell (sparse_sign.jl:339-348,
 * count_sketch.jl:166-213) for operands taller than one block. */
int rla_sparse_apply_bell_split_bytes(int64_t rows, int64_t in_rows, int64_t rb, int64_t n_rhs,
                                      size_t* bytes);
int rla_sparse_apply_bell_split_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                                    const int32_t* widths, const int32_t* bidx,
                                    const float* bval, int64_t n_rhs, const float* X,
                                    int64_t x_rs, int64_t x_cs, float* C, int64_t c_rs,
                                    int64_t c_cs, double alpha, double beta);
int rla_sparse_apply_bell_split_f64(rla_ctx* ctx, int64_t rows, int64_t in_rows, int64_t rb,
                                    const int32_t* widths, const int32_t* bidx,
                                    const double* bval, int64_t n_rhs, const double* X,
                                    int64_t x_rs, int64_t x_cs, double* C, int64_t c_rs,
                                    int64_t c_cs, double alpha, double beta);

/* Row-group tiles: the same left product for tall Float32 X when M (rows x
 * in_rows, CSR ptr/idx with strictly ascending columns per row) has at most
 * 4096 rows.  The lanes of a wavefront are 64 columns of X; a wavefront
 * keeps the FP32 sums of its output rows in registers (16 wavefronts, at
 * most 1024 rows, per workgroup; more rows take one workgroup per group of
 * 1024) while blocks of 384 rows of X pass through LDS.  Each entry of M is
 * read once for all 64 columns: with_values = 0 (nonzeros of one magnitude:
 * SparseSign, CountSketch) one 32-bit word, its LDS offset and its sign;
 * with_values = 1 (any values: FJLT's K) two, the offset and the value.
 * Each workgroup sums a range of blocks; the ranges' sums are added in
 * ascending order in FP64, so the result is within rounding of
 * rla_sparse_apply (not identical).  Replaces the same mul! as
 * rla_sparse_apply_bell (sparse_sign.jl:339-348, count_sketch.jl:166-213).
 *   rla_sparse_tile_shape: words per record and record count of the copy
 *     (RLA_ENOTSUP above 4096 rows or when a record would not fit the LDS,
 *     RLA_EDOMAIN for columns out of range or not strictly ascending);
 *   rla_sparse_tile_build_f32: fills the copy (16-byte aligned, nrec *
 *     rec_words words) and returns the magnitude (1 with values; RLA_EDOMAIN
 *     without values if the nonzeros do not share one magnitude);
 *   rla_sparse_apply_tile_f32: C = alpha*M*X + beta*C for X column-major
 *     (x_rs == 1); the context workspace holds the ranges' sums, at most
 *     rla_sparse_apply_tile_bytes bytes. */
int rla_sparse_tile_shape(rla_ctx* ctx, int64_t rows, int64_t in_rows, const int32_t* ptr,
                          const int32_t* idx, int with_values, int64_t* rec_words, int64_t* nrec);
int rla_sparse_tile_build_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, const int32_t* ptr,
                              const int32_t* idx, const float* val, int with_values,
                              int64_t rec_words, uint32_t* tiles, double* magnitude);
int rla_sparse_apply_tile_bytes(int64_t rows, int64_t in_rows, int64_t n_rhs, size_t* bytes);
int rla_sparse_apply_tile_f32(rla_ctx* ctx, int64_t rows, int64_t in_rows, int with_values,
                              int64_t rec_words, const uint32_t* tiles, double magnitude,
                              int64_t n_rhs, const float* X, int64_t x_rs, int64_t x_cs, float* C,
                              int64_t c_rs, int64_t c_cs, double alpha, double beta);

/* Agmon maximal-residual selection, the top-k step of sample_distribution!
 * for the Agmon distribution (src/Compressors/Distributions/agmon.jl:354-416):
 * out[t] (t < k) = the candidate with the t-th largest |v[c]|, ties to the
 * smaller index, over the candidates cand[0..n_cand) (device int32, 0-based;
 * NULL = 0..n_items-1 with n_cand = n_items).  v (length n_items) is the
 * residual A x - b (Left) or A'(A x - b) (Right).  A repeated candidate may be
 * chosen once per occurrence; a candidate outside 0..n_items-1 is never chosen
 * (out[t] = -1 when none is left).  1 <= k <= n_cand <= 2^20. */
int rla_agmon_select_f32(rla_ctx* ctx, int64_t n_items, const int32_t* cand, int64_t n_cand,
                         const float* v, int64_t k, int32_t* out);
int rla_agmon_select_f64(rla_ctx* ctx, int64_t n_items, const int32_t* cand, int64_t n_cand,
                         const double* v, int64_t k, int32_t* out);

/* ---- Walsh–Hadamard / SRHT / FJLT -------------------------------------
 * In-place FWHT of each of ncols columns of X (length n, a power of two),
 * natural (Sylvester) order, radix-2 stages h = 1, 2, 4, ... exactly as
 * fwht!(x, signs; scaling) / fwht!(x; scaling)
 * (src/Compressors/helpers/fwht.jl:21-118): the first stage multiplies each
 * entry by +-scaling (signs == NULL: +scaling).  Bit-identical to that loop. */
int rla_fwht_f32(rla_ctx* ctx, int64_t n, int64_t ncols, const uint64_t* sign_chunks,
                 double scaling, float* X, int64_t ldx);
int rla_fwht_f64(rla_ctx* ctx, int64_t n, int64_t ncols, const uint64_t* sign_chunks,
                 double scaling, double* X, int64_t ldx);

/* SRHT operator T = (scale * H * D * pad)[op, :] (s x m, H the n_pad-point
 * unnormalised Hadamard, D = diag(+-1) from sign_chunks, pad = zero-padding
 * of m rows to n_pad).
 *   adjoint = 0:  Y = alpha * T * X + beta * Y      (X m x ncols, Y s x ncols)
 *   adjoint = 1:  Y = alpha * T' * X + beta * Y     (X s x ncols, Y m x ncols)
 * Left S*A is adjoint 0 on A; Right A*S is adjoint 0 on A' (Y = C');
 * Left A*S is adjoint 1 on A'; Right S*A is adjoint 1 on A.
 * Replaces mul! at src/Compressors/srht.jl:220-504.  scale is the recipe's
 * stored Float64 (reference: type(sqrt(s)/sqrt(n_pad)), srht.jl:152).
 * plan: from rla_srht_plan for this op (NULL: built in the workspace on each
 * call).  beta == 0 does not read Y.
 * rla_srht_apply_f64 and rla_srht_apply_exact_f32 are bit-identical to the
 * reference loops for adjoint = 0: with unit row stride and 2^11 <= n_pad <=
 * 2^17 (Float32) or 2^16 (Float64) as one fused kernel, else multi-pass.
 * rla_srht_apply_f32 is the Float32 product within rounding of the reference
 * (1e-5 relative, the Float32 bar): adjoint = 0 with unit row stride,
 * 2^14 <= n_pad <= 2^24 and s <= 4096 runs the pruned kernel (only the
 * sampled rows of H D pad(X) are formed), other shapes as the exact entry. */
int rla_srht_apply_f32(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                       int64_t ncols, const int32_t* op, const uint64_t* sign_chunks,
                       double scale, const void* plan, const float* X, int64_t x_rs,
                       int64_t x_cs, float* Y, int64_t y_rs, int64_t y_cs, double alpha,
                       double beta);
int rla_srht_apply_exact_f32(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                             int64_t ncols, const int32_t* op, const uint64_t* sign_chunks,
                             double scale, const void* plan, const float* X, int64_t x_rs,
                             int64_t x_cs, float* Y, int64_t y_rs, int64_t y_cs, double alpha,
                             double beta);
int rla_srht_apply_f64(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                       int64_t ncols, const int32_t* op, const uint64_t* sign_chunks,
                       double scale, const void* plan, const double* X, int64_t x_rs,
                       int64_t x_cs, double* Y, int64_t y_rs, int64_t y_cs, double alpha,
                       double beta);

/* Sampled-row plan of an SRHT recipe (bitmap of needed positions, slot ->
 * output-row map) for the fused forward kernel; rebuild after the recipe's op
 * changes (update_compressor!).  plan_bytes is 0 when the shape has no fused
 * kernel (then pass plan = NULL). */
int rla_srht_plan_bytes(int elem_bytes, int64_t n_pad, int64_t s, size_t* bytes);
int rla_srht_plan(rla_ctx* ctx, int elem_bytes, int64_t n_pad, int64_t s, const int32_t* op,
                  const uint64_t* sign_chunks, void* plan);

/* ---- FJLT --------------------------------------------------------------
 * K = sprandn(type, s, n_pad, q) (src/Compressors/fjlt.jl:183, 225): each
 * entry independently nonzero with probability q (to within 2^-32), nonzero
 * values N(0,1), as an s x n_pad CSC matrix with rows ascending per column.
 * rla_fjlt_gen_count fills colptr[n_pad+1] and returns nnz to the host
 * (synchronises); rla_fjlt_gen_fill then writes rowidx[nnz] and val[nnz].
 * The index sets are a pure function of (seed, epoch) and identical to
 * rla_host_fjlt_gen_f64's; values agree with it to rounding.  The recipe's
 * signs come from rla_srht_gen with op = NULL. */
int rla_fjlt_gen_count(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad,
                       double q, int32_t* colptr, int64_t* nnz);
int rla_fjlt_gen_fill_f32(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad,
                          double q, const int32_t* colptr, int32_t* rowidx, float* val);
int rla_fjlt_gen_fill_f64(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad,
                          double q, const int32_t* colptr, int32_t* rowidx, double* val);
int rla_host_fjlt_gen_f64(uint64_t seed, uint64_t epoch, int64_t s, int64_t n_pad, double q,
                          int32_t* colptr_host, int32_t* rowidx_host, double* val_host,
                          int64_t cap, int64_t* nnz);

/* FJLT operator T = scale * K * H * D * pad (s x m; K above in CSC and its
 * CSR twin from rla_sparse_transpose; H, D, pad as for the SRHT).
 *   adjoint = 0:  Y = alpha * T * X + beta * Y      (X m x ncols, Y s x ncols)
 *   adjoint = 1:  Y = alpha * T' * X + beta * Y     (X s x ncols, Y m x ncols)
 * Left S*A is adjoint 0 on A; Right A*S is adjoint 0 on A' (Right recipes
 * hold K' of the reference's n_pad x s K); Left A*S is adjoint 1 on A';
 * Right S*A is adjoint 1 on A.  Replaces mul! at fjlt.jl:257-535; scale is the
 * recipe's type(1/(sqrt(n_pad) sqrt(q) sqrt(s))) (fjlt.jl:184, 226). */
int rla_fjlt_apply_f32(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                       int64_t ncols, const int32_t* colptr, const int32_t* rowidx,
                       const float* val, const int32_t* rowptr, const int32_t* colidx,
                       const float* val_t, const uint64_t* sign_chunks, double scale,
                       const float* X, int64_t x_rs, int64_t x_cs, float* Y, int64_t y_rs,
                       int64_t y_cs, double alpha, double beta);
int rla_fjlt_apply_f64(rla_ctx* ctx, int adjoint, int64_t m, int64_t n_pad, int64_t s,
                       int64_t ncols, const int32_t* colptr, const int32_t* rowidx,
                       const double* val, const int32_t* rowptr, const int32_t* colidx,
                       const double* val_t, const uint64_t* sign_chunks, double scale,
                       const double* X, int64_t x_rs, int64_t x_cs, double* Y, int64_t y_rs,
                       int64_t y_cs, double alpha, double beta);

/* The forward FJLT (adjoint = 0 above) with K's rows as the row-interleaved
 * copy from rla_sparse_ell_build (ptr = K's CSR row pointers): the gather's
 * lanes read consecutive words of K.  Bit-identical to rla_fjlt_apply with
 * adjoint = 0; replaces mul! at fjlt.jl:257-319 (and 395-468 for Right A*S). */
int rla_fjlt_apply_ell_f32(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                           const int32_t* ptr, const int32_t* ell_idx, const float* ell_val,
                           const uint64_t* sign_chunks, double scale, const float* X,
                           int64_t x_rs, int64_t x_cs, float* Y, int64_t y_rs, int64_t y_cs,
                           double alpha, double beta);
int rla_fjlt_apply_ell_f64(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                           const int32_t* ptr, const int32_t* ell_idx, const double* ell_val,
                           const uint64_t* sign_chunks, double scale, const double* X,
                           int64_t x_rs, int64_t x_cs, double* Y, int64_t y_rs, int64_t y_cs,
                           double alpha, double beta);

/* The forward FJLT product (T * X, fjlt.jl:257-319) with K (s x n_pad) given
 * by its block-interleaved copy (rla_sparse_bell_*, in_rows = n_pad, blocks of
 * rb rows).  Bit-identical to rla_fjlt_apply with adjoint = 0. */
int rla_fjlt_apply_bell_f32(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                            int64_t rb, const int32_t* widths, const int32_t* bidx,
                            const float* bval, const uint64_t* sign_chunks, double scale,
                            const float* X, int64_t x_rs, int64_t x_cs, float* Y, int64_t y_rs,
                            int64_t y_cs, double alpha, double beta);
int rla_fjlt_apply_bell_f64(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                            int64_t rb, const int32_t* widths, const int32_t* bidx,
                            const double* bval, const uint64_t* sign_chunks, double scale,
                            const double* X, int64_t x_rs, int64_t x_cs, double* Y, int64_t y_rs,
                            int64_t y_cs, double alpha, double beta);

/* The same forward FJLT product for Float32 with K given by its row-group
 * tile copy with values (rla_sparse_tile_* with with_values = 1, rows = s,
 * in_rows = n_pad; s <= 4096): the FWHT of each chunk of columns, then the
 * tile gather (its ranges' sums beside the transform in the context
 * workspace).  Within rounding of rla_fjlt_apply (fjlt.jl:257-319). */
int rla_fjlt_apply_tile_f32(rla_ctx* ctx, int64_t m, int64_t n_pad, int64_t s, int64_t ncols,
                            int64_t rec_words, const uint32_t* tiles,
                            const uint64_t* sign_chunks, double scale, const float* X,
                            int64_t x_rs, int64_t x_cs, float* Y, int64_t y_rs, int64_t y_cs,
                            double alpha, double beta);

/* ---- dense products and conversions ------------------------------------
 * Column-major GEMM C = alpha*op(A)*op(B) + beta*C (vendor library: the plain
 * Float32/Float64 products of the Gaussian compressor, gaussian.jl:145-170,
 * and of the plain-precision RangeFinder, power_its.jl:21-72). */
int rla_gemm_f32(rla_ctx* ctx, int trans_a, int trans_b, int64_t m, int64_t n, int64_t k,
                 double alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
                 double beta, float* C, int64_t ldc);
int rla_gemm_f64(rla_ctx* ctx, int trans_a, int trans_b, int64_t m, int64_t n, int64_t k,
                 double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                 double beta, double* C, int64_t ldc);

/* The same product on the FP64 MFMA (v_mfma_f64_16x16x4f64; split-K for
 * tall-skinny shapes, deterministic): the Gram and R^-1 products of
 * CholeskyQR2. */
int rla_gemm_f64_mfma(rla_ctx* ctx, int trans_a, int trans_b, int64_t m, int64_t n, int64_t k,
                      double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                      double beta, double* C, int64_t ldc);

/* Dense product of strided views on the FP32 / FP64 matrix cores
 * (v_mfma_f32_16x16x4_f32, v_mfma_f64_16x16x4_f64), products and sums in the
 * data's own precision: C = alpha * A * B + beta * C with A (m x k) at
 * A[i*a_rs + l*a_cs], B (k x n) at B[l*b_rs + j*b_cs], C at C[i*c_rs + j*c_cs].
 * A Julia column-major matrix is (p, 1, ld), its adjoint (p, ld, 1).  beta = 0
 * does not read C.  Deterministic: long-K products split K over workgroups and
 * add the partial sums in a fixed order.  Replaces the BLAS gemm!/gemv! calls of
 * the reference's dense paths (gaussian.jl:145-170, rangefinder.jl:151-191,
 * randsvd.jl:198-293, ihs.jl:220-249, col_projection.jl:282-334). */
int rla_gemm_strided_f32(rla_ctx* ctx, int64_t m, int64_t n, int64_t k, double alpha,
                         const float* A, int64_t a_rs, int64_t a_cs, const float* B, int64_t b_rs,
                         int64_t b_cs, double beta, float* C, int64_t c_rs, int64_t c_cs);
int rla_gemm_strided_f64(rla_ctx* ctx, int64_t m, int64_t n, int64_t k, double alpha,
                         const double* A, int64_t a_rs, int64_t a_cs, const double* B,
                         int64_t b_rs, int64_t b_cs, double beta, double* C, int64_t c_rs,
                         int64_t c_cs);

/* Gaussian compressor apply, gaussian.jl:145-170: side 0 (Left)
 * C = alpha * S * A + beta * C, side 1 (Right) C = alpha * A * S + beta * C,
 * with S = recipe.op (s_rows x s_cols, strided: pass the adjoint's strides for
 * S') and A (a_rows x a_cols, strided).  RLA_EDIM when the inner dimensions
 * differ.  The same kernels as rla_gemm_strided_*. */
int rla_gaussian_apply_f32(rla_ctx* ctx, int side, int64_t s_rows, int64_t s_cols,
                           const float* S, int64_t s_rs, int64_t s_cs, int64_t a_rows,
                           int64_t a_cols, const float* A, int64_t a_rs, int64_t a_cs,
                           double alpha, double beta, float* C, int64_t c_rs, int64_t c_cs);
int rla_gaussian_apply_f64(rla_ctx* ctx, int side, int64_t s_rows, int64_t s_cols,
                           const double* S, int64_t s_rs, int64_t s_cs, int64_t a_rows,
                           int64_t a_cols, const double* A, int64_t a_rs, int64_t a_cs,
                           double alpha, double beta, double* C, int64_t c_rs, int64_t c_cs);

/* Allocation-free rla_gaussian_apply_f32: no counterpart in the reference.  It
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results (bit for bit) as rla_gaussian_apply_f32, but
 * the split partial sums go to a file-scope device array of 32 MiB per device
 * (held until the process ends) instead of the context workspace; the split
 * rule keeps every shape within it (a larger need would fail with
 * RLA_ENOTSUP).  Calls that overlap on one device share that array and race. */
int rla_gaussian_apply_noalloc_f32(rla_ctx* ctx, int side, int64_t s_rows, int64_t s_cols,
                                   const float* S, int64_t s_rs, int64_t s_cs, int64_t a_rows,
                                   int64_t a_cols, const float* A, int64_t a_rs, int64_t a_cs,
                                   double alpha, double beta, float* C, int64_t c_rs,
                                   int64_t c_cs);

/* BF16 copies: dst[l + i*ldd] = bf16(src[i, l]) (round to nearest even) for
 * the rows x cols strided view src (rows become contiguous K-runs). */
int rla_to_bf16_f32(rla_ctx* ctx, int64_t rows, int64_t cols, const float* src, int64_t s_rs,
                    int64_t s_cs, uint16_t* dst, int64_t ldd);
int rla_to_bf16_f64(rla_ctx* ctx, int64_t rows, int64_t cols, const double* src, int64_t s_rs,
                    int64_t s_cs, uint16_t* dst, int64_t ldd);

/* BF16 MFMA GEMM, FP32 accumulation (v_mfma_f32_32x32x16_bf16):
 * C[i,j] = alpha * sum_l A[l + i*lda] * B[l + j*ldb] + beta * C[i*c_rs + j*c_cs],
 * i < M, j < N, l < K; A and B 16-byte aligned, lda and ldb multiples of 8.
 * The mixed-precision Gaussian sketch and power-iteration products
 * (power_its.jl:21,28-29,59,68,72). */
int rla_gemm_bf16(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, float alpha, const uint16_t* A,
                  int64_t lda, const uint16_t* B, int64_t ldb, float beta, float* C, int64_t c_rs,
                  int64_t c_cs);

/* Tiled BF16 copy of the rows x cols strided view src, the layout in which
 * the BF16 GEMM streams its A operand with one contiguous kilobyte per
 * wavefront load: tiles of 128 rows x 64 columns, tile (t, s) at element
 * (t * ceil(cols/64) + s) * 8192; inside a tile the 8 entries of row
 * 32 g + r (r < 32), columns 16 q + 8 h .. + 8 (h < 2), at element
 * ((4 g + q) * 64 + r + 32 h) * 8.  Entries past the view are written as
 * zero.  Same rounding as rla_to_bf16_*.  rla_to_bf16_tiled_bytes gives the
 * size of dst (16-byte aligned).  The resident copies of A for precision 2
 * of rla_rangefinder / rla_randsvd_finish (power_its.jl:21,28-29,59,68,72). */
int rla_to_bf16_tiled_bytes(int64_t rows, int64_t cols, size_t* bytes);
int rla_to_bf16_tiled_f32(rla_ctx* ctx, int64_t rows, int64_t cols, const float* src,
                          int64_t s_rs, int64_t s_cs, uint16_t* dst);
int rla_to_bf16_tiled_f64(rla_ctx* ctx, int64_t rows, int64_t cols, const double* src,
                          int64_t s_rs, int64_t s_cs, uint16_t* dst);
/* rla_gemm_bf16 with A (M x K) in the tiled layout above, N <= 288 (else
 * RLA_ENOTSUP): the same products in the same order, so the same bits. */
int rla_gemm_bf16_tiled(rla_ctx* ctx, int64_t M, int64_t N, int64_t K, float alpha,
                        const uint16_t* At, const uint16_t* B, int64_t ldb, float beta, float* C,
                        int64_t c_rs, int64_t c_cs);

int rla_convert_f32_f64(rla_ctx* ctx, int64_t rows, int64_t cols, const float* src, int64_t lds,
                        double* dst, int64_t ldd);
int rla_convert_f64_f32(rla_ctx* ctx, int64_t rows, int64_t cols, const double* src, int64_t lds,
                        float* dst, int64_t ldd);
/* dst = src * diag(d) (RandSVD's S scaling, randsvd.jl:224,270). */
int rla_scale_cols_f64(rla_ctx* ctx, int64_t rows, int64_t cols, const double* src, int64_t lds,
                       const double* d, double* dst, int64_t ldd);

/* ---- QR, RangeFinder, RandSVD -------------------------------------------
 * FP64 CholeskyQR2 of the m x k (m >= k) column-major Y in place (Y <- Q;
 * R optional k x k upper with Y_in = Q R), falling back to Householder QR
 * when the Cholesky factor breaks down; *used_householder reports which.
 * k <= 288: Gram matrix in one read of Y, Cholesky in one workgroup with
 * the trailing matrix in registers, Y*R^-1 by an in-place blocked solve (all
 * FP64 MFMA); wider k: tiled FP64 MFMA GEMM and the vendor solver's
 * potrf/trtri.
 * Replaces the qr! calls of power_its.jl:35,60,70,73. */
int rla_qr_workspace_bytes(int64_t m, int64_t k, size_t* bytes);
int rla_qr_f64(rla_ctx* ctx, int64_t m, int64_t k, double* Y, int64_t ldy, double* R,
               int64_t ldr, void* work, size_t work_bytes, int* used_householder);

/* x <- R^-1 x (trans = 0) or R'^-1 x (trans = 1) for an n x n upper
 * triangular R (column-major, ldr); x a device vector.  The two solves of the
 * IHS update, R'R u = g with R the triangular factor of S*A
 * (src/Solvers/ihs.jl:243-246). */
int rla_trsv_upper_f64(rla_ctx* ctx, int trans, int64_t n, const double* R, int64_t ldr,
                       double* x);

/* Cholesky factor and inverse of a k x k symmetric positive definite G
 * (k <= 288, else RLA_ENOTSUP; upper triangle used): R upper with G = R'R
 * (R may alias G), X = R^-1 (column-major, ldx; NULL to skip).  *info = 0,
 * or the 1-based column whose pivot was not above k * eps * G_jj (then R and
 * X are incomplete).
 * The factorisation step of the CholeskyQR that replaces power_its.jl's qr!. */
int rla_chol_inv_f64(rla_ctx* ctx, int64_t k, const double* G, int64_t ldg, double* R,
                     int64_t ldr, double* X, int64_t ldx, int* info);

/* RangeFinder iterations on Y = A*S (m x k, FP64, column-major, in/out):
 * orthogonalize != 0 runs rand_ortho_it (power_its.jl:53-80), otherwise
 * rand_power_it (power_its.jl:15-36); Y leaves as Q.
 * precision 0: products in A's type (a_elem 4 or 8; A column-major, lda) on
 * the vendor BLAS; A_rm, when not NULL, is a second copy of A in A's type
 * with rows contiguous (row stride ld_rm >= n), which turns every A'*X into
 * a plain (non-transposed) product; A_cm is unused.
 * precision 1: products on the BF16 MFMA with FP32 accumulation, reading the
 * two resident BF16 copies of A (A_rm: rows contiguous, ld_rm; A_cm:
 * columns contiguous, ld_cm; from rla_to_bf16_*).
 * precision 2: as 1 with the copies in the GEMM's tiled layout (A_rm =
 * rla_to_bf16_tiled of A, m x n; A_cm = rla_to_bf16_tiled of A', n x m;
 * ld_rm and ld_cm unused): the same products, bit-identical Q, with each
 * operand load one contiguous kilobyte.  QR is FP64 throughout. */
int rla_rangefinder_workspace_bytes(int precision, int a_elem, int64_t m, int64_t n, int64_t k,
                                    size_t* bytes);
int rla_rangefinder(rla_ctx* ctx, int precision, int orthogonalize, int64_t m, int64_t n,
                    int64_t k, int64_t power_its, int a_elem, const void* A, int64_t lda,
                    const void* A_rm, int64_t ld_rm, const void* A_cm, int64_t ld_cm,
                    double* Y, int64_t ldy, void* work, size_t work_bytes);

/* FP64 SVD of a k x k matrix, k <= 288 (the `svd(QA)` of randsvd.jl:186 once
 * QA is reduced to its k x k triangular factor): X = U diag(S) V' (X' when
 * transpose != 0), S descending, U and V k x k.  One-sided Jacobi on the
 * device (one cooperative launch); `sweeps` (host, may be NULL, then nothing
 * blocks) receives the number of Jacobi sweeps and makes the call wait for
 * the result.  A zero singular value gets a zero column of U. */
int rla_svd_small_workspace_bytes(int64_t k, size_t* bytes);
int rla_svd_small_f64(rla_ctx* ctx, int64_t k, const double* X, int64_t ldx, int transpose,
                      double* U, int64_t ldu, double* S, double* V, int64_t ldv, void* work,
                      size_t work_bytes, int* sweeps);

/* RandSVD finish (randsvd.jl:174-189) from the range Q (m x k): B' = A'Q,
 * B' = Qb Rb (CholeskyQR2), Rb' = Ur S Vr' (rla_svd_small_f64, k <= 288),
 * U = Q Ur, V = Qb Vr.  A_cm: for precision 1 the columns-contiguous BF16
 * copy of A, for precision 2 the tiled BF16 copy of A' (n x m); for
 * precision 0 NULL or the rows-contiguous copy of A in A's
 * type (as A_rm of rla_rangefinder).
 * Outputs U (m x k), S (k), V (n x k), FP64. */
int rla_randsvd_workspace_bytes(int precision, int a_elem, int64_t m, int64_t n, int64_t k,
                                size_t* bytes);
int rla_randsvd_finish(rla_ctx* ctx, int precision, int64_t m, int64_t n, int64_t k, int a_elem,
                       const void* A, int64_t lda, const void* A_cm, int64_t ld_cm,
                       const double* Q, int64_t ldq, double* U, int64_t ldu, double* S, double* V,
                       int64_t ldv, void* work, size_t work_bytes);

/* ---- Sampling compressor, Kaczmarz -------------------------------------
 * Row weights w_i = sum_j A[i,j]^2 in FP64 (L2Norm, strohmer_vershynin.jl:82;
 * column weights by passing transposed strides). */
int rla_row_norms2_f32(rla_ctx* ctx, int64_t m, int64_t n, const float* A, int64_t a_rs,
                       int64_t a_cs, double* w);
int rla_row_norms2_f64(rla_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t a_rs,
                       int64_t a_cs, double* w);

/* nsets weighted samples of k items from 0:n-1 (w NULL = uniform), set r
 * drawn with epoch+r, each sorted ascending: wsample!(state_space, weights,
 * idx; ordered=true, replace) of sample_distribution! (uniform.jl:93-95,
 * strohmer_vershynin.jl:150-153).  replace = 0: Efraimidis–Spirakis keys,
 * the s largest (successive sampling without replacement); replace = 1:
 * independent inverse-CDF draws.  Needs at least k items of positive weight
 * when replace = 0; 1 <= k <= 1024. */
int rla_wsample(rla_ctx* ctx, uint64_t seed, uint64_t epoch, int64_t n, const double* w,
                int64_t k, int replace, int64_t nsets, int32_t* sets);

/* ||b - A x||_2 accumulated in FP64 (FullResidual compute_error,
 * full_residual.jl:32-41); blocks until the value is on the host. */
int rla_residual_norm_f32(rla_ctx* ctx, int64_t m, int64_t n, const float* A, int64_t a_rs,
                          int64_t a_cs, const float* x, const float* b, double* out_host);

int rla_kaczmarz_workspace_bytes(int64_t m, int64_t n, int64_t s, int64_t every, size_t* bytes);

/* One block projection x += alpha * At' (At At')^-1 (bt - At x) with an
 * explicit s x n block At (strided) and bt (stride bt_inc): the LQ
 * minimum-norm step of kaczmarz_update_block! (kaczmarz.jl:294-311,
 * lq.jl:35-46); s = 1 is kaczmarz_update! (kaczmarz.jl:262-276).  A zero block
 * is skipped (kaczmarz.jl:300); a row dependent on the others gets a zero
 * coefficient (the projection onto the independent rows); a non-finite Gram
 * matrix returns RLA_EDOMAIN.  s <= 64.  Workspace from
 * rla_kaczmarz_workspace_bytes(s, n, s, 1). */
/* The same projection for any block size s <= n and Float32 / Float64 data,
 * as lq.jl:35-46 computes it: At' = Q R (FP64 CholeskyQR2, Householder on
 * breakdown), x += alpha Q (R' \\ (bt - At x)).  A zero block is skipped; rows
 * dependent on the others (a pivot of R not above sqrt(s eps) max|R_jj|) are
 * dropped and the block factored again, so the step projects onto the
 * independent rows.  Synchronises twice (the zero-block test and the pivot
 * check; once more when it drops rows).  Workspace from
 * rla_kaczmarz_lq_workspace_bytes(s, n). */
int rla_kaczmarz_lq_workspace_bytes(int64_t s, int64_t n, size_t* bytes);
int rla_kaczmarz_lq_update_f32(rla_ctx* ctx, int64_t s, int64_t n, const float* At,
                               int64_t at_rs, int64_t at_cs, const float* bt, int64_t bt_inc,
                               float* x, double alpha, void* work, size_t work_bytes);
int rla_kaczmarz_lq_update_f64(rla_ctx* ctx, int64_t s, int64_t n, const double* At,
                               int64_t at_rs, int64_t at_cs, const double* bt, int64_t bt_inc,
                               double* x, double alpha, void* work, size_t work_bytes);
/* ||b - A x||_2 in FP64 for Float64 data (full_residual.jl:32-41); blocks. */
int rla_residual_norm_f64(rla_ctx* ctx, int64_t m, int64_t n, const double* A, int64_t a_rs,
                          int64_t a_cs, const double* x, const double* b, double* out_host);
int rla_kaczmarz_block_update_f32(rla_ctx* ctx, int64_t s, int64_t n, const float* At,
                                  int64_t at_rs, int64_t at_cs, const float* bt, int64_t bt_inc,
                                  float* x, double alpha, void* work, size_t work_bytes);

/* rsolve! (kaczmarz.jl:313-368) with a Sampling compressor of s rows and
 * weights w (NULL: Uniform): for i = 1..max_it, the FP64 residual is
 * evaluated at i = 1, 1+every, 1+2*every, ... and at max_it, recorded in hist
 * and tested against threshold (BasicLogger threshold_stop; i == max_it always
 * stops, basic_logger.jl:93-98); otherwise a new row set is drawn and x is
 * projected.  A_rm: the matrix row-major (ld_rm), read by the block gathers
 * and the residual; A: the same matrix column-major (lda), only checked.
 * One launch per iteration (the previous block's update fused with the new
 * block's Gram matrix and solve).  x is updated in place.  s <= 64. */
int rla_kaczmarz_solve_f32(rla_ctx* ctx, int64_t m, int64_t n, const float* A, int64_t lda,
                           const float* A_rm, int64_t ld_rm, const float* b, float* x, int64_t s,
                           double alpha, const double* w, int replace, uint64_t seed,
                           uint64_t epoch0, int64_t max_it, int64_t every, double threshold,
                           double* hist, int64_t hist_cap, int64_t* n_hist, int64_t* iterations,
                           int* converged, void* work, size_t work_bytes);

/* ---- batched sketch-and-solve least squares ------------------------------
 * One problem: x = argmin || S A x - S b || with S the s x m CountSketch
 * given as CSR (s_rowptr[s+1], s_colidx (the rows of A in bucket r, ascending),
 * s_val (+-1)), i.e. CountSketch mul! (count_sketch.jl:166-213) followed by
 * the QRSolver solve x = qr!(SA) \ Sb (src/Solvers/SubSolvers/qr.jl:36-43).
 * A: m x n Float32 column-major (lda); b: m; x: n (written).  n <= 271,
 * n <= s <= 1024. */
typedef struct rla_ls_problem {
  int64_t m, n, lda;
  const float* A;
  const float* b;
  float* x;
  const int32_t* s_rowptr;
  const int32_t* s_colidx;
  const float* s_val;
  int64_t s;
} rla_ls_problem;

typedef struct rla_ls_plan rla_ls_plan;

/* Plans a batch: validates, allocates the per-problem state (SA, Gram,
 * factor) in one device allocation, picks the sketch tile shape (autotune: 0
 * default, 1 time the three candidates on this batch, 2..4 force candidate
 * 0..2, e.g. from a tile cache, and time it alone), and captures the batch (persistent
 * sketch kernel, batched FP64 Gram, batched Cholesky, back-substitution) in a
 * hipGraph.  The problem data must stay valid while the plan is used. */
int rla_ls_plan_create(rla_ctx* ctx, int64_t nprob, const rla_ls_problem* probs, int autotune,
                       rla_ls_plan** plan);
/* Replays the graph on the context's stream (asynchronous; no allocation). */
int rla_ls_plan_execute(rla_ls_plan* plan);
/* Per-problem status of the last execution: 0, or the 1-based column whose
 * pivot broke down (x not written).  Blocks. */
int rla_ls_plan_info(rla_ls_plan* plan, int32_t* info_host);
/* Chosen tile shape and the autotuner's timings (ms, -1 if not timed). */
int rla_ls_plan_config(rla_ls_plan* plan, int* cols_per_tile, int* rows_per_chunk,
                       float* tuned_ms3);
/* Copies problem p's sketched [SA Sb] (s x (n+1), column-major) to out. */
int rla_ls_plan_sketch(rla_ls_plan* plan, int64_t p, float* out);
int rla_ls_plan_destroy(rla_ls_plan* plan);

#ifdef __cplusplus
}
#endif

#endif /* RLA_MI355X_H */
