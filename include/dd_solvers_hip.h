/*
 * dd_solvers_hip.h — C ABI of the MI355X (gfx950) Schwarz-PCG kernels.
 *
 * Every entry point takes plain device pointers, 64-bit sizes, dtype codes and an explicit
 * hipStream_t (passed as void*), enqueues its kernels on that stream, and returns a
 * dds_status (0 on success). Arguments are validated on the host before anything is
 * launched; the message of the last failure on the calling thread is returned by
 * dds_last_error(). No entry point allocates device memory: callers (the Python package
 * uses the PyTorch caching allocator) pass every buffer in, which keeps the hot path legal
 * under hipGraph stream capture.
 *
 * Reference interfaces replaced (paths relative to the reference repository root):
 *   dds_gemv_strided_batched  dd_solvers_gemv::gemv_strided_batched
 *                             (src/dd_solvers/cpp/gemv.cpp:9-57, gemv_kernels.cu:31-153)
 *   dds_bsrmv                 FastBSR RPT-I numba kernel (src/dd_solvers/dd_solvers/bsr.py:40-76)
 *   dds_csrmv, dds_sellmv     torch sparse-CSR @ dense in CG(bsr_matmul=False) and
 *                             HybridSchwarz (src/dd_solvers/dd_solvers/solvers.py:226,235,264,880)
 *   dds_local_apply           Inv.solve -> gemv_strided_batched (solvers.py:355-359),
 *                             generalised to uneven and overlapping subdomains
 *   dds_local_apply_sym       the same on symmetric tile-packed inverses
 *   dds_local_apply_sym_mfma  the same for FP16/BF16 inverses on the matrix cores
 *   dds_sym_mfma_pack, dds_sym_mfma_apply_packed
 *                             its two kernels, launched apart
 *   dds_local_apply_sym_mfma_fused
 *                             the same for FP16 in one kernel (the pack in LDS)
 *   dds_local_apply_sym_mfma_noalloc
 *                             the same with its workspace in a static device array (no
 *                             counterpart in the reference)
 *   dds_local_apply_batched   the same on the reference's (n, k, k) layout (equal subdomains)
 *   dds_overlap_sum           y += repeat_interleave(...) prolongation (solvers.py:821-825,
 *                             :882-887) fused with the overlapped sum_i R_i^T y_i
 *   dds_restrict              x.reshape(n_s,-1).sum(1) + segment_reduce (solvers.py:817-820)
 *   dds_spd_inverse_batched, dds_inverse_batched
 *                             torch.linalg.inv in Inv.setup (solvers.py:350-352): SPD
 *                             (no pivoting) and general (partial pivoting) batches
 *   dds_chol_factor           cudssExecute(ANALYSIS|FACTORIZATION) of the coarse matrix
 *                             (src/dd_solvers/cpp/cudss.cpp:80-105), as a dense FP64 Cholesky
 *   dds_chol_solve            cudssExecute(SOLVE) (cudss.cpp:64-78)
 *   dds_chol_inverse, dds_dense_symv
 *                             coarse solve for large coarse spaces (explicit A_C^-1 from the
 *                             Cholesky factor, applied as one dense symmetric matvec)
 *   dds_nd_forward, dds_nd_backward, dds_nd_top_assemble, dds_nd_top_push
 *                             cudssExecute(SOLVE) of a large sparse coarse matrix
 *                             (cudss.cpp:64-78, called from CUDSS.solve, solvers.py:106-110):
 *                             the two sweeps of a nested-dissection block LDL^T factor
 *   dds_nd_plan_build, dds_nd_plan_field, dds_nd_plan_free
 *                             cudssExecute(ANALYSIS) in spdFactorize (cudss.cpp:80-96): the
 *                             ordering and symbolic tables, host code
 *   dds_nd_factor             cudssExecute(FACTORIZATION) (cudss.cpp:97-105): the fronts of
 *                             one tree height
 *   dds_mfma_apply, dds_mfma_apply_sub, dds_sym_mfma_apply16
 *                             the local apply to blocks of 16 right-hand sides (no counterpart:
 *                             the reference solves one right-hand side at a time)
 *   dds_pcg_*                 the CG.solve loop body (solvers.py:228-297): torch.linalg.norm,
 *                             torch.dot, fused_add / fused_sub / fused_mul_add (:170-189)
 *   dds_cast                  rhs.to(preconditioner_precision) / y.to(rhs.dtype)
 *                             (solvers.py:813, :826, :876, :888)
 */
#ifndef DD_SOLVERS_HIP_H
#define DD_SOLVERS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDS_ABI_VERSION 4
/* Upper bound on the per-block partial sums written by one reduction-producing launch
 * (partial buffers hold this many doubles, x16 for blocks of right-hand sides). */
#define DDS_MAX_PARTIALS 16384

typedef enum dds_dtype {
  DDS_F64 = 0,
  DDS_F32 = 1,
  DDS_F16 = 2,
  DDS_BF16 = 3,
  DDS_I32 = 4,
  DDS_I64 = 5
} dds_dtype;

typedef enum dds_status {
  DDS_OK = 0,
  DDS_EINVAL = 1,      /* bad argument (shape, dtype, pointer, size) */
  DDS_EHIP = 2,        /* HIP runtime error */
  DDS_EUNSUPPORTED = 3 /* valid but not implemented combination */
} dds_status;

/*
 * Device-resident PCG scalars (one per solve). The loop in CG.solve keeps alpha, beta and
 * rho on the device and reads ||r|| back every iteration (solvers.py:244); here the norm
 * test also runs on the device, so the host reads `done` only when it wants to.
 *   it      index of the next norm check (= number of completed updates)
 *   done    0 running, 1 converged (||r|| < atol at check `it`), 2 maxiter updates done
 * Gated kernels (those taking `const dds_pcg_state* gate`) return immediately when
 * gate->done != 0, so enqueuing iterations past convergence changes nothing.
 */
typedef struct dds_pcg_state {
  double atol;     /* max(atol, rtol * ||b||)                       solvers.py:229 */
  double rho;      /* r_i . z_i                                      :256 */
  double rho_prev; /* r_{i-1} . z_{i-1}                              :268 */
  double pq;       /* p_i . q_i                                      :265 */
  double alpha;    /* rho / pq                                       :265 */
  double beta;     /* rho / rho_prev                                 :259 */
  double rr;       /* ||r||^2 at the last check */
  int32_t it;
  int32_t done;
  int32_t maxiter;
  int32_t pad_;
} dds_pcg_state;

/* finalize ops (dds_pcg_finalize) */
#define DDS_FIN_CHECK 0 /* rr = sum; norms[it] = sqrt(rr); done = 1 if sqrt(rr) < atol */
#define DDS_FIN_RHO 1   /* rho = sum; beta = rho / rho_prev (it > 0) */
#define DDS_FIN_ALPHA 2 /* pq = sum; alpha = rho / pq; alphas[it] = alpha; betas[it] = beta */
#define DDS_FIN_NEXT 3  /* rho_prev = rho; it += 1; done = 2 if it == maxiter, else FIN_CHECK */

int dds_abi_version(void);
const char* dds_last_error(void);

/* ---- batched dense GEMV (drop-in op) ---------------------------------------------------
 * out[b, r] = sum_c mat[b, r, c] * vec[b, c] for b < n, with mat column-major
 * (element (b, r, c) at b*k*k + c*k + r), accumulated in the vector dtype.
 * mat_dtype in {F64, F32, F16, BF16}; vec_dtype in {F64, F32}; F64 matrices need F64
 * vectors (gemv.cpp:34-38). use_lds stages vec[b, :] in LDS (the reference's shared-memory
 * variant, gemv_kernels.cu:59-93). */
int dds_gemv_strided_batched(int mat_dtype, int vec_dtype, const void* mat, const void* vec,
                             void* out, int64_t n, int64_t k, int use_lds, void* stream);

/* ---- SpMV --------------------------------------------------------------------------------
 * BSR with b x b blocks stored column-major inside each block (element (r, c) of block j
 * at values[j*b*b + c*b + r], the cusparse csr2gebsr layout, cusparse.cpp:30, 81-83).
 * val_dtype in {F64, F32}; idx_dtype in {I32, I64} is the dtype of crow; col is int32.
 * If dot_partials != NULL the launch also writes per-block partial sums of x . y and
 * stores their count in *n_partials (host int); x is then read at the row indices too, so
 * the matrix must be square (len(x) >= rows). gate may be NULL. */
int dds_bsrmv(int val_dtype, int idx_dtype, int64_t n_block_rows, int64_t block_size,
              const void* crow, const int32_t* col, const void* values, const void* x, void* y,
              double* dot_partials, int* n_partials, const dds_pcg_state* gate, void* stream);

/* CSR y = A x; `lanes` lanes cooperate on one row (a power of two in [1, 1024]; the Python
 * layer picks the power of two nearest above nnz / n_rows). Up to 64 the lanes are part of
 * one wavefront; above 64 a workgroup of `lanes` threads handles each row (long rows, e.g.
 * R_C A of the hybrid coarse correction). Partials and gate as for dds_bsrmv. */
int dds_csrmv(int val_dtype, int idx_dtype, int64_t n_rows, const void* row_ptr,
              const int32_t* col, const void* values, const void* x, void* y, int lanes,
              double* dot_partials, int* n_partials, const dds_pcg_state* gate, void* stream);

/* SELL-64 y = A x: rows in slices of 64; entry j of row 64q + l at slot slice_ptr[q]/64 + j,
 * lane l: values[slice_ptr[q] + 64*j + l] (slice_ptr in elements, multiples of 64), padding
 * entries (col = own row, value 0). Column indices are compact: for slice q with
 * col_ptr[q] >= 0 the column is slot_base[slice_ptr[q]/64 + j] + col16[col_ptr[q] + 64*j + l]
 * (col16 read as unsigned 16-bit), so a slot costs 2 bytes per entry when its 64 columns
 * lie within a 65536-wide window; otherwise (col_ptr[q] < 0) it is
 * col32[~col_ptr[q] + 64*j + l]. The format CG uses for bsr_matmul=False (torch sparse-CSR
 * @ dense in the reference, solvers.py:280); built from CSR at setup. Partials and gate as
 * for dds_bsrmv. */
int dds_sellmv(int val_dtype, int64_t n_rows, const int64_t* slice_ptr, const int64_t* col_ptr,
               const uint16_t* col16, const int32_t* col32, const int32_t* slot_base,
               const void* values, const void* x, void* y, double* dot_partials,
               int* n_partials, const dds_pcg_state* gate, void* stream);

/* ---- local inverses --------------------------------------------------------------------
 * Y[sub_ptr[s] + a] = scale[s] * sum_b Ainv_s[a, b] * x[g_s(b)] for every subdomain s,
 * where g_s(b) = sub_idx[sub_ptr[s] + b] (or sub_ptr[s] + b when sub_idx == NULL) and
 * Ainv_s is stored column-major with leading dimension ld[s] at mats + mat_off[s]
 * (elements). ld[s] must be a multiple of dds_local_vec_width(mat_dtype) and >= k_s, the
 * padding rows must be finite, and mat_off[s] a multiple of the same width.
 * `tiles` lists (s, row0) pairs, row0 a multiple of dds_local_tile_rows(mat_dtype); one
 * workgroup per tile. vec_dtype (x and Y) in {F64, F32}; acc_dtype in {F64, F32};
 * scale may be NULL. max_k bounds every k_s (sizes the LDS stage of x). */
int dds_local_vec_width(int mat_dtype);
int dds_local_tile_rows(int mat_dtype);
int dds_local_apply(int mat_dtype, int vec_dtype, int acc_dtype, int64_t n_tiles,
                    const int32_t* tiles, const int64_t* sub_ptr, const int32_t* sub_idx,
                    int64_t max_k, const int64_t* mat_off, const int32_t* ld,
                    const float* scale, const void* mats, const void* x, void* Y,
                    int waves, const dds_pcg_state* gate, void* stream);

/* The local apply on the reference's own layout (Inv.solve -> gemv_strided_batched,
 * solvers.py:355-359; gemv_kernels.cu:31-153): n equal, contiguous, non-overlapping
 * subdomains of k unknowns, inverses (n, k, k) column-major; y[b*k + r] =
 * scale[b] * sum_c A_b[r][c] x[b*k + c] with the accumulator of the vector type (scale
 * optional), the drop-in GEMV kernels with a gate. use_lds: stage x in LDS (k <= 1024). */
int dds_local_apply_batched(int mat_dtype, int vec_dtype, const void* mats, const void* x,
                            void* y, int64_t n, int64_t k, int use_lds, const float* scale,
                            const dds_pcg_state* gate, void* stream);
/* Symmetric local inverses stored as the lower triangle of 64 x 64 tiles (layout in
 * csrc/local_sym.hip): each off-diagonal tile is read once and applied twice. One
 * wavefront per item (s, p) = (subdomain, pair of tile rows p and T_s-1-p) writes the
 * partial vector Yp[ypos[s] + p*64*T_s + a] for the first 64*(T_s - p) local rows a; the
 * partials are summed by dds_sym_combine. Replaces the same
 * Inv.solve -> gemv_strided_batched step (solvers.py:355-359) as dds_local_apply, with half
 * the matrix traffic. dtype rules as for dds_local_apply. */
int dds_local_apply_sym(int mat_dtype, int vec_dtype, int acc_dtype, int64_t n_items,
                        const int32_t* items, const int64_t* sub_ptr, const int32_t* sub_idx,
                        const int64_t* mat_off, const int64_t* ypos, const float* scale,
                        const void* mats, const void* x, void* Yp, const dds_pcg_state* gate,
                        void* stream);

/* The symmetric local apply for inverses stored in FP16 or BF16, on the matrix cores with
 * FP32 accumulation. Storage: the lower triangle of 64 x 64 tiles, each tile as its non-zero
 * 16 x 32 blocks in MFMA A-operand order (lane l of block (Rb, K2) holds
 * T[16Rb + l%16][32K2 + 8(l/16) + e], e < 8; diagonal tiles hold L + D/2). Items (s, q)
 * group pairs_per_item = G tile-row pairs: item q of subdomain s covers pairs
 * p = qG .. min(P_s, qG + G) - 1 (P_s = ceil(T_s / 2)), sums its products in FP32 and writes
 * Yp[ypos[s] + q*64*T_s + a] for the local rows a < min(k_s, 64 (T_s - qG)) once. When
 * G >= P_s for every subdomain each subdomain has one item and Yp = y, ypos = sub_ptr gives
 * the result directly; otherwise dds_sym_combine sums the items with the same G. Per apply
 * two kernels run: the first packs x (local row a of subdomain s is x[sub_idx[sub_ptr[s] + a]],
 * or x[sub_ptr[s] + a] when sub_idx is NULL) per 64-row block into device scratch: block b
 * (n_blk = blk_ptr[n_s] = sum_s ceil(k_s / 64) of them, blocks of subdomain s at
 * blk_ptr[s] .., blk_owner[b] = s) gets sig[b] (n_blk floats, the power of two that bounds
 * |x| on the block) and x / sig[b] split into 2 (FP16) or 3 (BF16) 16-bit terms in xt
 * (NT x 64 words per block); the second streams the tiles into LDS. vec_dtype F64 or F32;
 * max_k (the largest subdomain) at most 4096. Replaces the same
 * Inv.solve -> gemv_strided_batched step (solvers.py:355-359) as dds_local_apply,
 * accumulating in FP32 as gemv_kernels.cu:47-54 does for FP32 vectors. */
int dds_local_apply_sym_mfma(int mat_dtype, int vec_dtype, int64_t n_s, int64_t n_items,
                             int64_t max_k, int64_t pairs_per_item, const int32_t* items,
                             const int64_t* sub_ptr, const int32_t* sub_idx,
                             const int64_t* mat_off, const int64_t* ypos, const float* scale,
                             const void* mats, const void* x, int64_t n_blk,
                             const int32_t* blk_owner, const int64_t* blk_ptr, float* sig,
                             void* xt, void* Yp, const dds_pcg_state* gate, void* stream);

/* The two kernels of dds_local_apply_sym_mfma as separate launches (same arguments): the x
 * pack into sig / xt, and the apply from a packed x. A packed x stays valid until x changes,
 * so the apply can be launched (and timed) alone on one pack. */
int dds_sym_mfma_pack(int mat_dtype, int vec_dtype, int64_t n_s, int64_t n_blk,
                      const int32_t* blk_owner, const int64_t* blk_ptr, const int64_t* sub_ptr,
                      const int32_t* sub_idx, const void* x, float* sig, void* xt,
                      const dds_pcg_state* gate, void* stream);
int dds_sym_mfma_apply_packed(int mat_dtype, int vec_dtype, int64_t n_s, int64_t n_items,
                              int64_t max_k, int64_t pairs_per_item, const int32_t* items,
                              const int64_t* sub_ptr, const int64_t* blk_ptr,
                              const int64_t* mat_off, const int64_t* ypos, const float* scale,
                              const void* mats, const float* sig, const void* xt, void* Yp,
                              const dds_pcg_state* gate, void* stream);

/* dds_local_apply_sym_mfma for FP16 inverses with the x pack done by each workgroup in its
 * own LDS: one launch, no sig / xt workspace, the same result bits. x is vec_dtype (FP64 or
 * FP32), Yp is y_dtype (FP64, or FP32 under either x); sub_idx as above. Replaces the same
 * Inv.solve -> gemv_strided_batched step (solvers.py:355-359). */
int dds_local_apply_sym_mfma_fused(int mat_dtype, int vec_dtype, int y_dtype, int64_t n_s,
                                   int64_t n_items, int64_t max_k, int64_t pairs_per_item,
                                   const int32_t* items, const int64_t* sub_ptr,
                                   const int32_t* sub_idx, const int64_t* mat_off,
                                   const int64_t* ypos, const float* scale, const void* mats,
                                   const void* x, void* Yp, const dds_pcg_state* gate,
                                   void* stream);

/* dds_local_apply_sym_mfma without a caller workspace: sig and xt live in a static device
 * array of the library (32 MiB per GPU, held until the process ends; its address looked up
 * once per device on the first call there). No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). It allocates nothing. Limit: the workspace,
 * ceil(4 n_blk / 256) * 256 + n_blk * NT * 128 bytes (NT = 2 for FP16, 3 for BF16), must fit
 * in the 32 MiB (about 86,000 blocks of 64 rows in BF16, 129,000 in FP16); beyond it the call
 * fails with an error and launches nothing. Nothing in the array needs to start at zero (the
 * pack writes every word the apply reads). Every call on one GPU shares its array, so calls
 * that overlap on one GPU (different streams, or other threads) race. */
int dds_local_apply_sym_mfma_noalloc(int mat_dtype, int vec_dtype, int64_t n_s, int64_t n_items,
                                     int64_t max_k, int64_t pairs_per_item, const int32_t* items,
                                     const int64_t* sub_ptr, const int32_t* sub_idx,
                                     const int64_t* mat_off, const int64_t* ypos,
                                     const float* scale, const void* mats, const void* x,
                                     int64_t n_blk, const int32_t* blk_owner,
                                     const int64_t* blk_ptr, void* Yp,
                                     const dds_pcg_state* gate, void* stream);

/* y[sub_ptr[s] + a] = sum of the partial vectors that cover local row a of subdomain s, in
 * item order: items q = 0 .. min(Q_s - 1, (T_s-1-a/64) / G), Q_s = ceil(P_s / G), with
 * G = pairs_per_item (1 for dds_local_apply_sym): the result indexed like the output of
 * dds_local_apply. One lane per (s, a), coalesced; max_k <= 65535 * 256. */
int dds_sym_combine(int dtype, int64_t n_s, int64_t max_k, int64_t pairs_per_item,
                    const int64_t* sub_ptr, const int64_t* ypos, const void* Yp, void* y,
                    const dds_pcg_state* gate, void* stream);

/* z[j] = sum_{p in [dof_ptr[j], dof_ptr[j+1])} Y[dof_pos[p]]   (dof_ptr == NULL: z[j] = Y[j])
 *        + coarse_sign * yc[cell[j]]                           (yc != NULL)
 * in dtype (F64 or F32). Y is of y_dtype: dtype, or F32 under F64 vectors (the local results
 * of the matrix-core apply, accumulated in FP32), summed in dtype. If r != NULL (F64 only)
 * per-block partials of r . z are written. z may alias Y when dof_ptr == NULL and the two
 * dtypes agree. */
int dds_overlap_sum(int dtype, int64_t n, const int64_t* dof_ptr, const int32_t* dof_pos,
                    const void* Y, int y_dtype, const int32_t* cell, const void* yc,
                    double coarse_sign, void* z, const double* r, double* dot_partials,
                    int* n_partials, const dds_pcg_state* gate, void* stream);

/* xc[c] = sum_{j in [coarse_ptr[c], coarse_ptr[c+1])} x[j]: lanes = 0, one workgroup per
 * cell (large cells); lanes = 8, 16, 32 or 64, that many lanes per cell (small cells). */
int dds_restrict(int dtype, int64_t n_c, const int64_t* coarse_ptr, const void* x, void* xc,
                 int lanes, const dds_pcg_state* gate, void* stream);

/* ---- dense coarse solver ---------------------------------------------------------------
 * In-place Cholesky A = L L^T of an n x n row-major SPD matrix (lower triangle read and
 * overwritten by L; the strict upper triangle is left as is). *info (device int) is 0 on
 * success, else 1 + the index of the first non-positive pivot. F64 and F32. */
int dds_chol_factor(int dtype, int64_t n, void* a, int64_t lda, int* info, void* stream);
/* y = (L L^T)^{-1} x by forward then backward substitution in one workgroup. */
int dds_chol_solve(int dtype, int64_t n, const void* L, int64_t ldl, const void* x, void* y,
                   const dds_pcg_state* gate, void* stream);
/* Ainv = (L L^T)^{-1} as a full symmetric n x n row-major matrix. */
int dds_chol_inverse(int dtype, int64_t n, const void* L, int64_t ldl, void* ainv,
                     int64_t ldi, void* stream);
/* y = S x for a dense symmetric row-major S (one wavefront per row). */
int dds_dense_symv(int dtype, int64_t n, const void* s, int64_t lds, const void* x, void* y,
                   const dds_pcg_state* gate, void* stream);

/* ---- sparse coarse solver (nested-dissection block LDL^T) ------------------------------
 * The solve of A x = b with A = L D L^T from a multifrontal factorisation in a
 * nested-dissection numbering (perm[i] = original index of unknown i): one forward and one
 * backward launch per height of the dissection tree, the dense top front in between
 * (dds_nd_top_assemble, dds_dense_symv with its inverse, dds_nd_top_push). A front owns
 * unknowns [j0, j0+nJ), couples to nS unknowns of its ancestors, and stores the stacked
 * factor [F_JJ^-1 (nJ x nJ); W = F_SJ F_JJ^-1 (nS x nJ)] column-major at fac + foff (element
 * (r, c) at foff + c*(nJ+nS) + r). Between levels, a child writes its update vector into its
 * own slot of its parent's front, slots[wbase + maps[moff + s]] (entries no child maps stay
 * zero), and a parent writes x at a child's coupling set into that child's buffer,
 * xs[push_dst[k]] = x at front position push_src[k]. Work items are 12-int64 records, one
 * per workgroup; lds_elems (<= 7680) is the largest nJ (forward) or nS (backward) of the
 * launch. dtype is that of the vectors and buffers (b, g, slots, xs, out), fac_dtype that of
 * the factor: F64/F64, F32/F32, or F64 vectors with an F32 factor (a factor stored in FP32,
 * for plans without a top front). Sums in FP64 in a fixed order. No allocation; b and out
 * are in the original numbering.
 *
 * Forward record (j0, nJ, nS, foff, soff, n_children, wbase, moff, r0, nr, R, -), R in
 * {64, 128, 256} and nr <= R (the 256 threads are R rows x 256/R column parts): with
 * f_J = b[perm[j0 + .]] + sum over the children q of slots[soff + q*(nJ+nS) + .], row
 * r of [r0, r0+nr): r < nJ: g[j0 + r] = (F_JJ^-1 f_J)_r; r >= nJ: slots[wbase +
 * maps[moff + r - nJ]] = (children's slots at r) - (W f_J)_{r - nJ}. */
int dds_nd_forward(int dtype, int fac_dtype, int64_t n_items, const int64_t* items,
                   const int32_t* perm, const int32_t* maps, const void* fac, const void* b,
                   void* g, void* slots, int64_t lds_elems, const dds_pcg_state* gate,
                   void* stream);
/* Backward record (j0, nJ, nS, wtoff, xoff, c0, nc, k0, k1, C, -, -), C in {16, 32, 64}
 * and nc <= C (the 256 threads are C columns x 256/C row parts), W row-major at
 * fac + wtoff: x_J[c] = g[j0 + c] - (W^T x_S)_c for c in [c0, c0+nc) with
 * x_S = xs[xoff .. xoff+nS), out[perm[j0 + c]] =
 * x_J[c]; then push entries [k0, k1) (sources in [c0, c0+nc), or >= nJ for the coupling
 * set). */
int dds_nd_backward(int dtype, int fac_dtype, int64_t n_items, const int64_t* items,
                    const int32_t* perm, const int64_t* push_dst, const int32_t* push_src,
                    const void* fac, const void* g, void* xs, void* out, int64_t lds_elems,
                    const dds_pcg_state* gate, void* stream);
/* Top front, unknowns [j0, j0+n_top): ftop[p] = b[perm[j0 + p]] + sum over its n_children
 * children q of slots[soff + q*n_top + p]. */
int dds_nd_top_assemble(int dtype, int64_t j0, int64_t n_top, int64_t soff, int64_t n_children,
                        const int32_t* perm, const void* b, const void* slots, void* ftop,
                        const dds_pcg_state* gate, void* stream);
/* out[perm[j0 + p]] = xtop[p] for p < n_top; xs[push_dst[i]] = xtop[push_src[i]] for
 * i < n_push. */
int dds_nd_top_push(int dtype, int64_t j0, int64_t n_top, int64_t n_push, const int32_t* perm,
                    const void* xtop, const int64_t* push_dst, const int32_t* push_src,
                    void* xs, void* out, const dds_pcg_state* gate, void* stream);

/* ---- PCG vector kernels (F64) ----------------------------------------------------------
 * Each writes per-block partial sums where stated (count returned in *n_partials). */
/* r = b - q; partials of r . r */
int dds_pcg_residual(int64_t n, const double* b, const double* q, double* r,
                     double* partials, int* n_partials, const dds_pcg_state* gate,
                     void* stream);
/* p = z (st->it == 0) else p = st->beta * p + z */
int dds_pcg_update_p(int64_t n, const double* z, double* p, const dds_pcg_state* st,
                     void* stream);
/* x += alpha p; r -= alpha q; partials of r . r */
int dds_pcg_update_xr(int64_t n, const double* p, const double* q, double* x, double* r,
                      double* partials, int* n_partials, const dds_pcg_state* st,
                      void* stream);
/* partials of a . b */
int dds_dot_partials(int64_t n, const double* a, const double* b, double* partials,
                     int* n_partials, const dds_pcg_state* gate, void* stream);
/* One-workgroup reduction of `count` partials (fixed order) and the scalar update `op`
 * (DDS_FIN_*). norms/alphas/betas hold maxiter + 1 doubles. */
int dds_pcg_finalize(int op, const double* partials, int count, dds_pcg_state* st,
                     double* norms, double* alphas, double* betas, void* stream);
/* Plain sum of `count` partials into *out (device double), for dot products outside PCG. */
/* Symbolic analysis (host code, no device work): the nested-dissection ordering of the
 * symmetric pattern (n x n CSR, int64 row pointers, int32 column indices sorted and unique
 * within each row; the diagonal may be present) and every table of the factorisation and of
 * the sweeps above, held by the library until dds_nd_plan_free. Tree nodes are split down to
 * leaf_size unknowns; nodes of depth < d_top, the largest bound whose nodes hold at most
 * top_max unknowns together, form the dense top front. Fails (DDS_EINVAL) when a lower
 * front exceeds the sweeps' 7680-unknown LDS capacity. Replaces the analysis phase of
 * cudssExecute(ANALYSIS) in spdFactorize (src/dd_solvers/cpp/cudss.cpp:80-96), which orders
 * the coarse matrix on the host as well. */
typedef struct dds_nd_plan dds_nd_plan;
int dds_nd_plan_build(int64_t n, const int64_t* indptr, const int32_t* indices,
                      int64_t leaf_size, int64_t top_max, dds_nd_plan** plan);
/* Fields of a plan (element type; length). F fronts, L heights (the top front's height
 * last), L' = L minus the top's height. dds_nd_plan_field copies a field into dst (unless
 * dst is NULL) and returns its length in elements, -1 for an unknown field. */
enum dds_nd_field {
  DDS_ND_SCALARS = 0,   /* int64; DDS_ND_N_SCALARS, indexed by dds_nd_scalar */
  DDS_ND_PERM = 1,      /* int64; n: perm[i] = original index of unknown i */
  DDS_ND_FJ0 = 2,       /* int64; F: first unknown of each front */
  DDS_ND_FNJ = 3,       /* int64; F: nJ */
  DDS_ND_FNS = 4,       /* int64; F: nS */
  DDS_ND_LEVEL_PTR = 5, /* int64; L + 1: fronts of height h are [level_ptr[h], level_ptr[h+1]) */
  DDS_ND_FOFF = 6,      /* int64; F + 1: stacked factor offsets */
  DDS_ND_WTOFF = 7,     /* int64; F + 1: row-major W offsets (wtoff[F] = factor size) */
  DDS_ND_SOFF = 8,      /* int64; F + 1: forward slot offsets */
  DDS_ND_WBASE = 9,     /* int64; F: the front's slot in its parent's front (-1: none) */
  DDS_ND_MOFF = 10,     /* int64; F + 1: coupling-set offsets (maps, s_idx, backward buffers) */
  DDS_ND_MAPS = 11,     /* int32; moff[F]: positions of a coupling set in the parent front */
  DDS_ND_PUSH_DST = 12, /* int64; push entries */
  DDS_ND_PUSH_SRC = 13, /* int32; push entries */
  DDS_ND_PUSH_PTR = 14, /* int64; F + 1 */
  DDS_ND_PARENT = 15,   /* int64; F (-1: none) */
  DDS_ND_CHILD_PTR = 16,/* int64; F + 1 */
  DDS_ND_CHILD_IDX = 17,/* int64; children of each front, in slot order */
  DDS_ND_S_IDX = 18,    /* int64; moff[F]: coupling sets (permuted numbering, ascending) */
  DDS_ND_AP_PTR = 19,   /* int64; n + 1: the permuted pattern A[perm][:, perm] */
  DDS_ND_AP_IDX = 20,   /* int32; nnz: its column indices, ascending per row */
  DDS_ND_AP_SRC = 21,   /* int64; nnz: position of each of its entries in the input */
  DDS_ND_FWD_PTR = 22,  /* int64; L' + 1: forward items per height */
  DDS_ND_FWD_ITEMS = 23,/* int64; 12 per forward item (dds_nd_forward records) */
  DDS_ND_FWD_LDS = 24,  /* int64; L': lds_elems of each forward launch */
  DDS_ND_BWD_PTR = 25,  /* int64; L' + 1 */
  DDS_ND_BWD_ITEMS = 26,/* int64; 12 per backward item (dds_nd_backward records) */
  DDS_ND_BWD_LDS = 27,  /* int64; L' */
  DDS_ND_UOFF2 = 28     /* int64; F + 1: update-matrix offsets (nS x nS per front) */
};
enum dds_nd_scalar {
  DDS_ND_N = 0, DDS_ND_FRONTS = 1, DDS_ND_TOP = 2 /* top-front unknowns (0: none) */,
  DDS_ND_DTOP = 3, DDS_ND_LEVELS = 4, DDS_ND_LOWER_LEVELS = 5, DDS_ND_FAC_SIZE = 6,
  DDS_ND_SLOT_SIZE = 7, DDS_ND_XS_SIZE = 8, DDS_ND_NNZ = 9, DDS_ND_TREE_NODES = 10,
  DDS_ND_N_SCALARS = 11
};
int64_t dds_nd_plan_field(const dds_nd_plan* plan, int field, void* dst);
void dds_nd_plan_free(dds_nd_plan* plan);
/* Numeric factorisation of the fronts [lo, hi) of one height (FP64, setup), one workgroup
 * per front. Front t's matrix F (M x M, M = nJ + nS, row-major at work + work_off[t - lo]) is
 * assembled from the entries of the permuted matrix in its columns (ap_ptr/ap_idx/ap_val;
 * rows J, columns >= j0, coupling columns found in s_idx[moff[t] ..]) and the update
 * matrices of its children (ustore + uoff2[c], nS_c x nS_c row-major, added at maps[moff[c]
 * ..] child by child in slot order). mode 0: the first nJ pivots of F are eliminated by
 * blocked Gauss-Jordan (no pivoting), which leaves F_JJ^-1, -W = -F_SJ F_JJ^-1 and the
 * Schur complement U = F_SS - W F_JS in place; the stacked factor [F_JJ^-1; W] goes to fac +
 * foff[t] (column-major), W row-major to fac + wtoff[t], U to ustore + uoff2[t]; info[t - lo]
 * = 1 when a pivot is not positive (not SPD). mode 1 (the top front): assembly only, the
 * dense front is then factored by dds_chol_factor and inverted by dds_chol_inverse into fac
 * + foff[t]. Replaces cudssExecute(FACTORIZATION) in spdFactorize (cudss.cpp:97-105). */
int dds_nd_factor(int mode, int64_t lo, int64_t hi, const int64_t* fj0, const int64_t* fnj,
                  const int64_t* fns, const int64_t* foff, const int64_t* wtoff,
                  const int64_t* moff, const int64_t* uoff2, const int64_t* child_ptr,
                  const int64_t* child_idx, const int32_t* maps, const int64_t* s_idx,
                  const int64_t* ap_ptr, const int32_t* ap_idx, const double* ap_val,
                  const int64_t* work_off, double* work, double* ustore, double* fac,
                  int32_t* info, void* stream);

int dds_sum_partials(const double* partials, int count, double* out, void* stream);

/* ---- setup -------------------------------------------------------------------------------
 * In-place inverse of n_mats SPD matrices (dtype DDS_F64 or DDS_F32): matrix b is row-major
 * k[b] x k[b] at A + b * stride elements, leading dimension ld (k[b] <= ld, stride >=
 * ld * k[b]); blocked Gauss-Jordan without pivoting, one workgroup per matrix. info[b]
 * (optional) is set to 1 when a pivot is not positive (the matrix is not SPD; the result is
 * then unusable). Replaces torch.linalg.inv(matrix_batch) in Inv.setup (solvers.py:350-352)
 * for the local matrices of the Schwarz preconditioners, which are SPD. */
int dds_spd_inverse_batched(int dtype, int64_t n_mats, const int32_t* k, int64_t ld,
                            int64_t stride, void* A, int32_t* info, void* stream);
/* In-place inverse of n_mats general matrices, laid out as for dds_spd_inverse_batched
 * (ld <= 4096, stride >= ld * ld): Gauss-Jordan elimination with partial pivoting by rows,
 * one workgroup per matrix. info[b] (optional) is set to q + 1 when column q has no nonzero
 * pivot (the matrix is singular; the result is then unusable). Replaces
 * torch.linalg.inv(matrix_batch) in Inv.setup (solvers.py:350-352) for any nonsingular
 * batch, and serves the Schwarz setup's local matrices that are not SPD. */
int dds_inverse_batched(int dtype, int64_t n_mats, const int32_t* k, int64_t ld, int64_t stride,
                        void* A, int32_t* info, void* stream);
/* Segmented sum (dtype DDS_F64 or DDS_F32): out[i] = sum of vals[seg_ptr[i] ..
 * seg_ptr[i+1]) added in index order in dtype, one thread per segment; compensated != 0
 * carries each addition's rounding error (Neumaier), for long runs that cancel. The
 * duplicate merge of sorted (key, value) lists at setup: the coarse matrix R_C A R_C^T
 * (solvers.py:721-771) and R0A = R_C A (solvers.py:890-906), compensated, and gather_csr
 * (gather_csr_sum_kernel, cusparse.py:78-158, which adds each run in order), plain. */
int dds_segment_sum(int dtype, int compensated, int64_t n_seg, const int64_t* seg_ptr,
                    const void* vals, void* out, void* stream);

/* ---- measurement ---------------------------------------------------------------------------
 * Streaming bandwidth over n doubles (n even, 16-byte aligned arrays), grid workgroups of 256
 * (grid <= DDS_MAX_PARTIALS): op 0 = read (y[b] = sum of a over workgroup b's elements; y holds
 * grid doubles), op 1 = copy (y = a), op 2 = triad (y = a + s * b). The achievable-bandwidth
 * reference for the memory-bound kernels; it replaces the clone-bandwidth measurement of
 * experiments/memory_bandwidth.ipynb (cells 2-4). */
int dds_stream(int op, int64_t n, const double* a, const double* b, double s, double* y,
               int64_t grid, void* stream);

/* ---- differentiation -------------------------------------------------------------------
 * grad[e] = -(1/N) exp(theta[e]) lam_e^T Khat u_e for the N^3 Q1 elements of the unit cube:
 * the coefficient gradient of J(u(theta)) given the adjoint solution lam of A lam = dJ/du.
 * lam, u hold the (N-1)^3 interior unknowns; map[i] (optional) is the storage position of
 * natural (x-fastest) interior node i. No counterpart in the reference (it has no autograd). */
int dds_q1_element_grad(int64_t N, const double* theta, const double* lam, const double* u,
                        const int32_t* map, double* grad, void* stream);

/* ---- blocks of 16 right-hand sides --------------------------------------------------------
 * Vectors are (n, 16) row-major FP64; st points to 16 dds_pcg_state (one per column, NULL
 * for ungated use); partials are laid out [block][16] (count returned in *n_partials).
 * Update kernels skip converged columns; gated kernels return when all 16 are done.
 * They replace the same reference steps as their single-vector versions above. */
int dds_pcg_residual_m(int64_t n, const double* b, const double* q, double* r, double* partials,
                       int* n_partials, void* stream);
int dds_pcg_update_p_m(int64_t n, const double* z, double* p, const dds_pcg_state* st,
                       void* stream);
int dds_pcg_update_xr_m(int64_t n, const double* p, const double* q, double* x, double* r,
                        double* partials, int* n_partials, const dds_pcg_state* st, void* stream);
int dds_dot_m(int64_t n, const double* a, const double* b, double* partials, int* n_partials,
              const dds_pcg_state* st, void* stream);
int dds_pcg_finalize_m(int op, const double* partials, int count, dds_pcg_state* st,
                       double* norms, double* alphas, double* betas, void* stream);
/* SELL-64 (slices of dds_sellmv, with plain int32 columns col[slice_ptr[q] + 64*j + l]) times
 * 16 vectors; optional x . y partials (square only). */
int dds_sellmm(int64_t n_rows, const int64_t* slice_ptr, const int32_t* col, const double* values,
               const double* x, double* y, double* partials, int* n_partials,
               const dds_pcg_state* st, void* stream);
/* CSR with long rows (R0A) times 16 vectors, one wavefront per row. */
int dds_csrmm_rows(int64_t n_rows, const int64_t* row_ptr, const int32_t* col,
                   const double* values, const double* x, double* y, const dds_pcg_state* st,
                   void* stream);
/* Y = S X, S dense symmetric n x n row-major (the coarse inverse). */
int dds_dense_symm_m(int64_t n, const double* S, const double* x, double* y,
                     const dds_pcg_state* st, void* stream);
int dds_restrict_m(int64_t n_c, const int64_t* coarse_ptr, const double* x, double* xc,
                   const dds_pcg_state* st, void* stream);
int dds_overlap_sum_m(int64_t n, const int64_t* dof_ptr, const int32_t* dof_pos, const double* Y,
                      const int32_t* cell, const double* yc, double sign, double* z,
                      const double* r, double* partials, int* n_partials,
                      const dds_pcg_state* st, void* stream);
/* Local inverses on MFMA (layouts in csrc/multi.hip): dds_mfma_pack_x scales X_s per
 * (subdomain, column) by its largest magnitude (sigma, n_s x 16 floats) and splits it into
 * 2 FP16 or 3 BF16 terms in B-operand order; dds_mfma_apply computes
 * Y[sub_ptr[s] + a][:] = scale[s] * sigma[s][:] * (A_s^-1 X_s)[a][:] with
 * v_mfma_f32_16x16x32 (FP32 accumulation), one wavefront per item (s, 64-row group),
 * `grid` persistent workgroups of `waves` wavefronts (grid 0: one per item group). Replace
 * Inv.solve -> gemv_strided_batched (solvers.py:355-359) for stacked right-hand sides. */
int dds_mfma_pack_x(int mat_dtype, int64_t n_s, const int64_t* sub_ptr, const int32_t* sub_idx,
                    const int64_t* b_off, const double* x, void* bp, float* sigma,
                    const dds_pcg_state* st, void* stream);
int dds_mfma_apply(int mat_dtype, int64_t n_items, const int32_t* items, const int64_t* sub_ptr,
                   const int64_t* a_off, const int64_t* b_off, const float* scale,
                   const float* sigma, const void* A, const void* bp, double* Y, int grid,
                   int waves, const dds_pcg_state* st, void* stream);
/* The same product, one workgroup of `waves` (4, 8) wavefronts per subdomain order[b]
 * (b < n_s; the order of the work, longest first): its waves take the groups of rpw (2, 4, 8)
 * row blocks in turn, so a subdomain's packed right-hand sides are fetched from HBM once and
 * shared through the L2 of its CU's XCD. Same bits as dds_mfma_apply for every shape. */
int dds_mfma_apply_sub(int mat_dtype, int64_t n_s, const int32_t* order, const int64_t* sub_ptr,
                       const int64_t* a_off, const int64_t* b_off, const float* scale,
                       const float* sigma, const void* A, const void* bp, double* Y, int waves,
                       int rpw, const dds_pcg_state* st, void* stream);
/* The same product from the matrix-core symmetric store of dds_local_apply_sym_mfma (lower
 * 64 x 64 tiles, half the bytes of the full inverses), for max_k <= 1024: one workgroup of 4
 * wavefronts per subdomain order[b] reads each tile once for Y_I += T X_J and
 * Y_J += T^T X_I. dds_sym_mfma_pack16 scales X per 64-row block b (blk_ptr / blk_owner as in
 * dds_sym_mfma_pack) and column: sigma[b][c] a power of two above max |X_b[:, c]| (n_blk x 16
 * floats), X_b / sigma[b] split into 2 FP16 or 3 BF16 terms (bp: n_blk x 2 x NT x 512
 * elements); dds_sym_mfma_apply16 writes Y[sub_ptr[s] + a][:] = scale[s] (A_s^-1 X_s)[a][:]. */
int dds_sym_mfma_pack16(int mat_dtype, int64_t n_blk, const int32_t* blk_owner,
                        const int64_t* blk_ptr, const int64_t* sub_ptr, const int32_t* sub_idx,
                        const double* x, float* sigma, void* bp, const dds_pcg_state* st,
                        void* stream);
int dds_sym_mfma_apply16(int mat_dtype, int64_t n_s, int64_t max_k, const int32_t* order,
                         const int64_t* sub_ptr, const int64_t* blk_ptr, const int64_t* mat_off,
                         const float* scale, const float* sigma, const void* mats,
                         const void* bp, double* Y, const dds_pcg_state* st, void* stream);

/* dst = (dst_dtype) src, F64 <-> F32 <-> F64 */
int dds_cast(int src_dtype, int dst_dtype, int64_t n, const void* src, void* dst,
             const dds_pcg_state* gate, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DD_SOLVERS_HIP_H */
