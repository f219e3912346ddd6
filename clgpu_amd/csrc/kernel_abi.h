/*
 * kernel_abi.h -- the thin C ABI between the C++ host layer (handle, validation,
 * implementation registry, memory staging) and the gfx950 kernels.
 *
 * Everything here works on device pointers and a HIP stream and returns a
 * hipError_t; nothing allocates. This replaces the reference's engine layer
 * (include/iclgpu/engine.hpp:72-423, core/src/ocl/): kernels are compiled ahead
 * of time for gfx950 and looked up in static tables instead of being JIT-built
 * from a string database.
 */
#pragma once
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct icl_gemm_args {
    int M, N, K;
    const float* A; /* complex kernels: interleaved {re, im} */
    const float* B;
    float* C;
    long lda, ldb, ldc; /* in elements of the data type */
    float alpha_re, alpha_im, beta_re, beta_im;
    int mask; /* output triangle mask, ICL_MASK_* */
    /* Level-3 extensions, honoured by the FP32 MFMA kernels only (all zero for a
     * plain GEMM): */
    int write_tri; /* ICL_TRI_*: store only this triangle of C, leave the rest untouched */
    int kband;     /* ICL_KBAND_*: per-tile k window of a zero-padded triangular operand */
    int cflags;    /* ICL_CFLAG_*: complex epilogue options */
} icl_gemm_args;

enum {
    ICL_TRI_NONE = 0,
    ICL_TRI_LOWER = 1, /* m >= n */
    ICL_TRI_UPPER = 2, /* m <= n */
    /* Flags on LOWER / UPPER, splitting a square triangle at its diagonal blocks
     * of ICL_TRI_BAND_BLOCK: OFF_BAND launches only the tiles outside them (wholly
     * inside the triangle), BAND only the tiles inside them; the two launches
     * write disjoint parts that together make the triangle. Tile sizes must
     * divide ICL_TRI_BAND_BLOCK. */
    ICL_TRI_OFF_BAND = 4,
    ICL_TRI_BAND = 8,
    ICL_TRI_BAND_BLOCK = 256
};

/* k window of a tile at rows [m0, m0 + BM), columns [n0, n0 + BN): */
enum {
    ICL_KBAND_NONE = 0,
    ICL_KBAND_K_BELOW_M_END = 1, /* op(A) lower triangular: k < m0 + BM */
    ICL_KBAND_K_FROM_M0 = 2,     /* op(A) upper triangular: k >= m0 */
    ICL_KBAND_K_BELOW_N_END = 3, /* op(B) upper triangular: k < n0 + BN */
    ICL_KBAND_K_FROM_N0 = 4      /* op(B) lower triangular: k >= n0 */
};

enum {
    ICL_CFLAG_CONJ_ACC = 1, /* C = alpha * conj(op(A) op(B)) + beta * C */
    ICL_CFLAG_REAL_DIAG = 2 /* imaginary part of C(m, m) stored as 0 */
};

enum {
    ICL_MASK_NONE = 0,
    ICL_MASK_LOWER = 1,        /* keep m >= n, zero elsewhere */
    ICL_MASK_LOWER_STRICT = 2, /* keep m >  n */
    ICL_MASK_UPPER = 3,        /* keep m <= n */
    ICL_MASK_UPPER_STRICT = 4  /* keep m <  n */
};

enum { ICL_KIND_SGEMM = 0, ICL_KIND_CGEMM = 1 };
enum { ICL_MATH_FP32 = 0, ICL_MATH_BF16X3 = 1 };

/* One compiled GEMM implementation (a tile configuration with all its
 * op/edge instantiations). opa/opb: 0 = N, 1 = T, 2 = C. edge = 0 requires
 * M % bm == N % bn == K % bk == 0 and 16-byte aligned operand columns.
 * workspace (may be NULL): device bytes the launch needs in `ws` (e.g. for
 * operands re-packed by a prepass); launch receives ws (NULL if none needed). */
typedef struct icl_gemm_variant {
    const char* name;
    int kind;
    int math;
    int bm, bn, bk;
    int threads;
    hipError_t (*launch)(const icl_gemm_args* args, int opa, int opb, int edge, void* ws, hipStream_t stream);
    size_t (*workspace)(const icl_gemm_args* args);
    float tflops_hint; /* measured useful TFLOP/s of the main kernel (a prepass is timed apart)
                          on a full MI355X (m = n = k = 8192, NN) */
    int occupancy;     /* workgroups resident per CU */
    int level3;        /* implements the level-3 fields write_tri, kband, cflags */
} icl_gemm_variant;

/* Kernel timing hook. When the calling thread's handle has kernel timing on,
 * records a HIP event on `stream`; otherwise does nothing. The host layer marks
 * the launch of a variant without a workspace; a variant with a prepass marks
 * its main kernel itself (right before and right after that launch), so the
 * prepass stays outside the timed pair. */
hipError_t icl_timer_mark(hipStream_t stream);

/* A second stream of the handle whose call is running, with two events, on
 * which a GEMM may run its operand prepass beside its first products (the
 * packed split-BF16 GEMMs); NULL when the call has none. */
typedef struct icl_side_lane {
    hipStream_t stream;
    hipEvent_t start, done;
} icl_side_lane;
const icl_side_lane* icl_gemm_side_lane(void);

/* Registered GEMM implementations. Kernel translation units register their
 * tables at library load (icl_register_gemm_variants); the host layer reads
 * them back with icl_gemm_variant_count / icl_gemm_variant_at, ordered by name. */
void icl_register_gemm_variants(const icl_gemm_variant* table, int count);
int icl_gemm_variant_count(void);
const icl_gemm_variant* icl_gemm_variant_at(int index);

/* Triangular solve of one diagonal block (at most 64 x 64) in canonical
 * forward-substitution form L * X = s_pre * X0 (L lower triangular), with
 * L(i, k) = L[i * l_si + k * l_sk] and X(i, j) = X[i * x_si + j * x_sj];
 * strides may be negative (backward substitution and right-side solves are
 * expressed by reversing / transposing the index maps). Each of the nlines
 * columns j is independent. The solution is scaled by s_post on store. */
typedef struct icl_trsm_diag_args {
    int nb;
    int nlines;
    const float* L;
    long l_si, l_sk;
    float* X;
    long x_si, x_sj;
    int unit;
    float s_pre, s_post;
} icl_trsm_diag_args;
hipError_t icl_trsm_diag(const icl_trsm_diag_args* a, hipStream_t stream);

/* Complex version of the diagonal-block solve: L and X hold interleaved
 * {re, im} pairs and the strides count complex elements. conj: every element
 * of L read is conjugated (op = conjugate transpose). X = s_post * (L^-1 X0),
 * with the division and products of the reference's complex helpers. */
typedef struct icl_ctrsm_diag_args {
    int nb;
    int nlines;
    const float* L;
    long l_si, l_sk;
    float* X;
    long x_si, x_sj;
    int unit;
    int conj;
    float s_post_re, s_post_im;
} icl_ctrsm_diag_args;
hipError_t icl_ctrsm_diag(const icl_ctrsm_diag_args* a, hipStream_t stream);

/* Inverts the 64 x 64 diagonal blocks of the canonical complex system with nsys
 * unknowns (L(i, k) = L[i * l_si + k * l_sk], interleaved complex, conjugated
 * when conj) into invD: ceil(nsys/64) dense column-major 64 x 64 complex blocks. */
typedef struct icl_ctrsm_blocked_args {
    int nsys;
    const float* L;
    long l_si, l_sk;
    int unit;
    int conj;
    float* invD;
} icl_ctrsm_blocked_args;
hipError_t icl_ctrsm_invert_blocks(const icl_ctrsm_blocked_args* a, hipStream_t stream);

/* X = W * B for one diagonal panel of a blocked TRSM: X(r, l) = sum_k W(r, k)
 * B(k, l) for r, k < p (p <= 256) and l < nlines, with W(r, k) = W[r * w_rs +
 * k * w_ks], B(k, l) = B[k * b_ks + l * b_ls], X(r, l) = X[r * x_rs + l * x_ls]
 * (elements; interleaved complex when cplx). One of w_rs, w_ks, one of b_ks,
 * b_ls and one of x_rs, x_ls should be 1 for coalesced access. */
typedef struct icl_panel_apply_args {
    int cplx, p, nlines;
    const float* W;
    long w_rs, w_ks;
    const float* B;
    long b_ks, b_ls;
    float* X;
    long x_rs, x_ls;
} icl_panel_apply_args;
hipError_t icl_panel_apply(const icl_panel_apply_args* a, hipStream_t stream);

/* Grouped products of depth 256 for the panel solves of the blocked TRSMs: up to
 * ICL_PANEL_GEMM_MAX problems in one launch, each
 *   D(i, l) = alpha * sum_{k < 256} P(i, k) Q(k, l) + beta * D(i, l)
 * for i < rows, l < lines (D is not read when beta == 0), with
 * P(i, k) = P[i * p_is + k * p_ks], Q(k, l) = Q[k * q_ks + l * q_ls] and
 * D(i, l) = D[i * d_is + l * d_ls] in elements. Every problem of a launch must
 * have the same contiguous dimensions: p_is == 1 or p_ks == 1, q_ks == 1 or
 * q_ls == 1, d_is == 1 or d_ls == 1, alike in all problems. No D may overlap a
 * P or Q of the launch. tile0 is set by icl_panel_gemm. */
#define ICL_PANEL_GEMM_MAX 16
typedef struct icl_panel_gemm_problem {
    int rows, lines;
    const float* P;
    long p_is, p_ks;
    const float* Q;
    long q_ks, q_ls;
    float* D;
    long d_is, d_ls;
    float alpha, beta;
    int tile0;
    int conj_p, conj_q; /* complex launches only: use conj(P) / conj(Q) */
} icl_panel_gemm_problem;
typedef struct icl_panel_gemm_args {
    int nprob;
    icl_panel_gemm_problem prob[ICL_PANEL_GEMM_MAX];
} icl_panel_gemm_args;
hipError_t icl_panel_gemm(icl_panel_gemm_args* a, hipStream_t stream);
/* The same for interleaved complex matrices (element (i, k) of P at float offset
 * 2 (i p_is + k p_ks), likewise Q and D), real alpha and beta, conj_p / conj_q
 * conjugating P / Q. Only the LDS-DMA kernel exists: every operand must allow
 * 16-byte copies (16-byte aligned, non-unit stride even) and a rows-contiguous
 * operand have an even row count; otherwise hipErrorInvalidValue and nothing is
 * launched (icl_cpanel_gemm_ok tells beforehand). */
hipError_t icl_cpanel_gemm(icl_panel_gemm_args* a, hipStream_t stream);
int icl_cpanel_gemm_ok(const icl_panel_gemm_args* a);

/* Inverses of the 256 x 256 diagonal panels of the canonical complex system
 * (panel p = unknowns [256 p, 256 p + plen), plen = min(256, nsys - 256 p)),
 * one workgroup per 64-wide column block of a panel, from the inverted 64 x 64
 * blocks invD
 * (icl_ctrsm_invert_blocks): M_jj = inv(L_jj) and, for i > j,
 * M_ij = -inv(L_ii) * sum_{k=j}^{i-1} L_ik M_kj. Element (r, c), r, c < plen, of
 * panel p's inverse is stored (interleaved complex) at element
 * p * 65536 + (rev ? (plen-1-r) * w_rs + (plen-1-c) * w_cs : r * w_rs + c * w_cs)
 * of W; entries above the block diagonal are stored as 0. */
typedef struct icl_ctrsm_panel_inv_args {
    int nsys;
    const float* L;
    long l_si, l_sk;
    int conj;
    const float* invD;
    float* W;
    long w_rs, w_cs;
    int rev;
    float* T; /* scratch: 4 blocks of 64 x 64 complex per panel */
} icl_ctrsm_panel_inv_args;
hipError_t icl_ctrsm_invert_panels(const icl_ctrsm_panel_inv_args* a, hipStream_t stream);

/* X_b = s_post * inv(L_bb) X_b in place for one diagonal block of nb <= 64
 * unknowns and nlines lines: X(i, j) = X[i * x_si + j * x_sj] (interleaved
 * complex), invD = that block's dense inverse. */
typedef struct icl_ctrsm_apply_args {
    int nb;
    int nlines;
    const float* invD;
    float* X;
    long x_si, x_sj;
    float s_post_re, s_post_im;
} icl_ctrsm_apply_args;
hipError_t icl_ctrsm_apply(const icl_ctrsm_apply_args* a, hipStream_t stream);

/* Blocked solve of the canonical lower-triangular system L X = B with nsys
 * unknowns and nlines right-hand sides: L(i, k) = L[i * l_si + k * l_sk],
 * B(i, j) = B[i * x_si + j * x_sj] (strides may be negative). The 64 x 64
 * diagonal blocks are inverted into invD (ceil(nsys/64) * 4096 floats); step b
 * forms X_b = inv(L_bb) B_b into X (X(i, j) = X[i * xv_si + j * xv_sj]) and
 * subtracts L_rb X_b from the trailing rows of B (B is used as scratch);
 * scatter writes alpha * X back through the view of B. The panel solve works on
 * rows [ps, pe) of the same views (see icl_trsm_panel). */
typedef struct icl_trsm_blocked_args {
    int nsys, nlines;
    const float* L;
    long l_si, l_sk;
    float* B;
    long x_si, x_sj;
    float* X;
    long xv_si, xv_sj;
    float* invD;
    int unit;
    int* sync;       /* dataflow solve only: icl_trsm_dataflow_sync_ints(...) ints, zeroed by invert_blocks */
    int* error_word; /* dataflow solve only: set to 1 (never cleared by the kernels) when a wait is aborted */
    int wait_polls;  /* dataflow solve only: polls of a flag per wait before the wait aborts; 0 = the
                      * default bound (65536), < 0 = abort the first wait without polling */
} icl_trsm_blocked_args;
hipError_t icl_trsm_invert_blocks(const icl_trsm_blocked_args* a, hipStream_t stream);
hipError_t icl_trsm_step(const icl_trsm_blocked_args* a, int block, hipStream_t stream);
hipError_t icl_trsm_scatter(const icl_trsm_blocked_args* a, float alpha, hipStream_t stream);

/* The whole blocked solve in one launch. X is tile-major here: the 64 x 64
 * tile (row block j, line block c) holds X(64 j + k, 64 c + l) at
 * X[(j * ncb + c) * 4096 + k * 64 + l], ncb = ceil(nlines / 64); xv_si and
 * xv_sj are unused. One workgroup per tile takes tiles in row-block order from a ticket
 * counter, accumulates L_bj X_jc over j < b as the producers' flags are set,
 * then X_bc = inv(L_bb) (B_bc - sum), stores it, sets its flag and writes
 * alpha * X_bc through the view of B. invert_blocks (same stream, just before)
 * zeroes a->sync (per-tile flags, the ticket counter, then the abort word). A
 * wait that exceeds a->wait_polls polls sets the abort word (the last int), after
 * which every wait returns at once, so the launch always drains; the result is
 * then invalid, and *error_word is set for the host to report. */
int icl_trsm_dataflow_sync_ints(int nsys, int nlines);

/* Solve of the diagonal panel rows [ps, pe) (pe - ps <= 256, ps a multiple of
 * 64) of the canonical system, for all lines: one workgroup per 64 lines runs
 * the panel's 64-row blocks in order inside the workgroup (X_i = inv(L_ii) B_i,
 * then B_k -= L_ki X_i for the later blocks k of the panel) and writes X rows
 * [ps, pe) through the X view. B rows of the panel are read, not written; rows
 * outside the panel are not touched (the caller updates them, e.g. with a GEMM). */
hipError_t icl_trsm_panel(const icl_trsm_blocked_args* a, int ps, int pe, hipStream_t stream);

/* Inverses W_p of the npanels full 256 x 256 diagonal panels of the canonical
 * lower-triangular L (L(i, k) = L[i * l_si + k * l_sk]) from its inverted 64 x 64
 * diagonal blocks invD (icl_trsm_invert_blocks; block b at invD + 4096 b,
 * element (r, c) at [c * 64 + r]): panel p's inverse element (r, c) goes to
 * W[p * 65536 + r' * w_rs + c' * w_cs] with (r', c') = (r, c), or
 * (255 - r, 255 - c) when rev. T is scratch of npanels * 16384 floats. */
typedef struct icl_trsm_panel_inv_args {
    int npanels;
    const float* L;
    long l_si, l_sk;
    const float* invD;
    float* W;
    long w_rs, w_cs;
    int rev;
    float* T;
} icl_trsm_panel_inv_args;
hipError_t icl_trsm_invert_panels(const icl_trsm_panel_inv_args* a, hipStream_t stream);
hipError_t icl_trsm_dataflow(const icl_trsm_blocked_args* a, float alpha, hipStream_t stream);

/* R = fp32(alpha * B0 - P * Q) with the products and the sum in FP64.
 * P is M x K, Q is K x N (fp32, column-major, optionally transposed); the
 * operand selected by tri_on (1 = P, 2 = Q) is triangular: entries outside
 * the triangle (lower if tri_lower) are taken as 0 and, if tri_unit, the
 * diagonal as 1. */
typedef struct icl_residual_args {
    int M, N, K;
    const float* P;
    long ldp;
    int p_trans;
    const float* Q;
    long ldq;
    int q_trans;
    const float* B0;
    long ldb0;
    float* R;
    long ldr;
    double alpha;
    int tri_on;
    int tri_lower;
    int tri_unit;
} icl_residual_args;
hipError_t icl_residual_f64(const icl_residual_args* a, hipStream_t stream);

/* Refinement convergence test: partial[2 b] = max |D| and partial[2 b + 1] =
 * max |X| over block b's share of the m x n matrices D and X (column-major,
 * leading dimensions ldd, ldx), for ICL_ABSMAX_BLOCKS blocks; the host takes the
 * maxima of the partials. NaNs are ignored. */
#define ICL_ABSMAX_BLOCKS 1024
hipError_t icl_absmax_pair(int m, int n, const float* D, long ldd, const float* X, long ldx, float* partial,
                           hipStream_t stream);

/* Y(m x n) = alpha * X + beta * Y elementwise (beta == 0: Y not read;
 * X may be NULL when alpha == 0). */
hipError_t icl_axpby_2d(int m, int n, float alpha, const float* X, long ldx, float beta, float* Y, long ldy,
                        hipStream_t stream);
/* Complex version of the same (used for CGEMM with k == 0 or alpha == 0). */
hipError_t icl_caxpby_2d(int m, int n, float alpha_re, float alpha_im, const float* X, long ldx, float beta_re,
                         float beta_im, float* Y, long ldy, hipStream_t stream);

/* y = alpha * op(A) * x + beta * y (op 0 = N, 1 = T, 2 = conjugate transpose for
 * complex), A m x n column-major; y is not read when beta == 0. icl_sgemv works on
 * floats, icl_cgemv on interleaved complex (lda, incx, incy in elements of the
 * type). splits > 1 splits the reduction dimension over the grid; partial then
 * holds splits * len(y) elements (len(y) = m for op N, n otherwise), summed by a
 * second kernel in split order. icl_gemv_splits gives the split count the
 * launch is sized for. incx, incy >= 1. */
typedef struct icl_gemv_args {
    int trans;
    int m, n;
    const float* A;
    long lda;
    const float* x;
    long incx;
    float* y;
    long incy;
    float alpha_re, alpha_im, beta_re, beta_im;
    float* partial;
    int splits;
} icl_gemv_args;
int icl_gemv_splits(int cplx, int trans, int m, int n);
hipError_t icl_sgemv(const icl_gemv_args* a, hipStream_t stream);
hipError_t icl_cgemv(const icl_gemv_args* a, hipStream_t stream);

/* Writes the dense order x order matrix W (column-major, leading dimension ldw)
 * that a triangular or symmetric operand stands for, so that the level-3
 * routines can run it through the GEMM kernels. A holds one triangle (lower if
 * `lower`), the other triangle of A is never read.
 *   ICL_EXPAND_TRI: W = op(A) with op 0 = N, 1 = T, 2 = conjugate transpose;
 *                   entries outside the triangle are 0 and, if unit, the
 *                   diagonal is 1 (the diagonal of A is then not read).
 *   ICL_EXPAND_SYM: W(i, j) = A(i, j) inside the stored triangle and A(j, i)
 *                   (conjugated if herm) outside it; herm also stores the
 *                   diagonal's imaginary parts as 0.
 * Complex data (icl_expand_complex) is interleaved {re, im}; lda/ldw count
 * elements of the data type. */
enum { ICL_EXPAND_TRI = 0, ICL_EXPAND_SYM = 1 };
typedef struct icl_expand_args {
    int order;
    const float* A;
    long lda;
    float* W;
    long ldw;
    int mode;
    int lower;
    int trans;
    int unit;
    int herm;
} icl_expand_args;
hipError_t icl_expand_real(const icl_expand_args* a, hipStream_t stream);
hipError_t icl_expand_complex(const icl_expand_args* a, hipStream_t stream);

/* The library's statically allocated device scratch (ICLBLAS_NOALLOC_SCRATCH_BYTES
 * per device, iclblasSgemmNoAlloc's workspace). icl_noalloc_scratch_init resolves
 * its address on `device`, which must be current; once per device, later calls
 * return the first result. icl_noalloc_scratch returns that address, or NULL
 * before a successful init on that device. */
hipError_t icl_noalloc_scratch_init(int device);
void* icl_noalloc_scratch(int device);


#ifdef __cplusplus
} /* extern "C" */
#endif
