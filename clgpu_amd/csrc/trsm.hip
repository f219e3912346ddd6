// STRSM building blocks for gfx950: the diagonal-block solver, the FP64
// residual used by iterative refinement, and 2-D elementwise updates.
//
// The substitution follows the reference kernel's per-column recurrence
// (libs/blas/funcs/Strsm/Strsm_naive.cl:28-37): for pivot i, x_i = b_i / a_ii
// (non-unit), then b_k -= a_ki * x_i for every later k, scaling by alpha last.
#include "mfma.hpp"
#include "kernel_abi.h"

namespace {

constexpr int kNB = 64; // largest diagonal block; one lane owns one right-hand side

// A lower-triangular diagonal block (nb <= 64) staged in LDS transposed,
// Lt[i][k] = L(k, i) for k > i, so that the update of pivot i reads a contiguous
// run of the column below the diagonal with every lane on the same address
// (broadcast). Unused entries are 0 so the fully unrolled update never reads
// garbage; diag[i] = L(i, i) (1 for a unit diagonal, which is never read).
struct DiagBlock {
    float Lt[kNB][kNB + 4];
    float diag[kNB];
};

__device__ __forceinline__ void stage_diag_block(DiagBlock& s, const float* L, long l_si, long l_sk, int nb, int unit,
                                                 int lane)
{
    for (int i = 0; i < kNB; ++i) {
        const int k = lane;
        float v = 0.0f;
        if (i < nb && k < nb && k > i) v = L[(long)k * l_si + (long)i * l_sk];
        s.Lt[i][k] = v;
    }
    if (lane < nb) s.diag[lane] = unit ? 1.0f : L[(long)lane * l_si + (long)lane * l_sk];
}

// Forward substitution L x = x in registers (the per-column recurrence of the
// reference kernel, Strsm_naive.cl:28-37).
__device__ __forceinline__ void solve_in_registers(float (&x)[kNB], const DiagBlock& s, int nb, int unit)
{
#pragma unroll
    for (int i = 0; i < kNB; ++i) {
        if (i < nb) {
            float xi = x[i];
            if (!unit) xi = xi / s.diag[i];
            x[i] = xi;
#pragma unroll
            for (int k = i + 1; k < kNB; ++k) x[k] -= s.Lt[i][k] * xi;
        }
    }
}

// One wavefront solves L * x = s_pre * b for 64 independent right-hand sides,
// one per lane, with x held in 64 registers.
__global__ void __launch_bounds__(64) trsm_diag_kernel(icl_trsm_diag_args a)
{
    __shared__ DiagBlock s;
    const int lane = threadIdx.x;
    const int nb = a.nb;
    stage_diag_block(s, a.L, a.l_si, a.l_sk, nb, a.unit, lane);
    __syncthreads();

    const long j = (long)blockIdx.x * kNB + lane;
    const bool active = j < a.nlines;
    float* xcol = a.X + j * a.x_sj;

    float x[kNB];
#pragma unroll
    for (int i = 0; i < kNB; ++i) x[i] = (active && i < nb) ? a.s_pre * xcol[(long)i * a.x_si] : 0.0f;
    solve_in_registers(x, s, nb, a.unit);
    if (active) {
#pragma unroll
        for (int i = 0; i < kNB; ++i)
            if (i < nb) xcol[(long)i * a.x_si] = a.s_post * x[i];
    }
}

// Inverts every 64 x 64 diagonal block of the canonical lower-triangular system
// (workgroup b handles rows/columns [64 b, 64 b + nb)): lane j solves
// L_bb y = e_j in registers, and column j of inv(L_bb) is written to
// invD[b * 4096 + j * 64 + i] (column-major, rows and columns beyond nb zero).
__global__ void __launch_bounds__(64) trsm_invert_kernel(icl_trsm_blocked_args a)
{
    __shared__ DiagBlock s;
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const int r0 = b * kNB;
    const int nb = min(kNB, a.nsys - r0);
    stage_diag_block(s, a.L + (long)r0 * (a.l_si + a.l_sk), a.l_si, a.l_sk, nb, a.unit, lane);
    __syncthreads();
    float x[kNB];
#pragma unroll
    for (int i = 0; i < kNB; ++i) x[i] = (i == lane && lane < nb) ? 1.0f : 0.0f;
    solve_in_registers(x, s, nb, a.unit);
    float* col = a.invD + (long)b * kNB * kNB + (long)lane * kNB;
#pragma unroll
    for (int i = 0; i < kNB; ++i) col[i] = (lane < nb && i < nb) ? x[i] : 0.0f;
}

// ---------------------------------------------------------------------------
// FP64 residual R = fp32(alpha * B0 - P * Q) on v_mfma_f64_16x16x4_f64.
// 64 x 64 output tile per workgroup (4 waves, each 32 x 32 = 2 x 2 MFMA blocks),
// K slabs of 32 staged in LDS as FP64; slabs outside the triangle are skipped.
// The next slab is read into registers while the current one is multiplied.
// ---------------------------------------------------------------------------
constexpr int RT = 64, RK = 32;
constexpr int kResPer = RT * RK / 256; // elements of each operand slab per thread

__device__ __forceinline__ float tri_load(const float* X, long ld, int trans, int r, int c, int rows, int cols,
                                          bool tri, int lower, int unit)
{
    // element (r, c) of op(X) (rows x cols), with the triangle rule applied
    if (r >= rows || c >= cols) return 0.0f;
    if (tri) {
        if (r == c && unit) return 1.0f;
        if (lower ? (c > r) : (c < r)) return 0.0f;
    }
    return trans ? X[(long)r * ld + c] : X[(long)c * ld + r];
}

__global__ void __launch_bounds__(256) residual_f64_kernel(icl_residual_args a)
{
    __shared__ double Ps[RK][RT + 1];
    __shared__ double Qs[RK][RT + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_m = (a.M + RT - 1) / RT;
    const int m0 = (blockIdx.x % tiles_m) * RT;
    const int n0 = (blockIdx.x / tiles_m) * RT;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;

    // K range that can be non-zero for this tile.
    int k_lo = 0, k_hi = a.K;
    if (a.tri_on == 1) { // P = op(A) is M x K triangular; rows m0..m0+63
        if (a.tri_lower) k_hi = min(a.K, m0 + RT);
        else k_lo = m0;
    } else if (a.tri_on == 2) { // Q = op(A) is K x N triangular; cols n0..n0+63
        if (a.tri_lower) k_lo = n0;
        else k_hi = min(a.K, n0 + RT);
    }
    k_lo = (k_lo / RK) * RK;

    f64x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

    // P slab: RT rows x RK k; Q slab: RK k x RT cols; the lanes run along the
    // contiguous dimension of each operand.
    auto p_coords = [&](int e, int& r, int& k) {
        if (a.p_trans) { k = e % RK; r = e / RK; } else { r = e % RT; k = e / RT; }
    };
    auto q_coords = [&](int e, int& k, int& c) {
        if (a.q_trans) { c = e % RT; k = e / RT; } else { k = e % RK; c = e / RK; }
    };
    float pv[kResPer], qv[kResPer];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int s = 0; s < kResPer; ++s) {
            int r, k, c;
            p_coords(tid + 256 * s, r, k);
            pv[s] = tri_load(a.P, a.ldp, a.p_trans, m0 + r, k0 + k, a.M, a.K, a.tri_on == 1, a.tri_lower,
                             a.tri_unit);
            q_coords(tid + 256 * s, k, c);
            qv[s] = tri_load(a.Q, a.ldq, a.q_trans, k0 + k, n0 + c, a.K, a.N, a.tri_on == 2, a.tri_lower,
                             a.tri_unit);
        }
    };
    if (k_lo < k_hi) fetch(k_lo);
    for (int k0 = k_lo; k0 < k_hi; k0 += RK) {
#pragma unroll
        for (int s = 0; s < kResPer; ++s) {
            int r, k, c;
            p_coords(tid + 256 * s, r, k);
            Ps[k][r] = (double)pv[s];
            q_coords(tid + 256 * s, k, c);
            Qs[k][c] = (double)qv[s];
        }
        __syncthreads();
        if (k0 + RK < k_hi) fetch(k0 + RK); // in flight during the products below
#pragma unroll
        for (int kk = 0; kk < RK; kk += 4) {
            const int kr = kk + (lane >> 4);
            double p[2], q[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) p[i] = Ps[kr][wm + 16 * i + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) q[j] = Qs[kr][wn + 16 * j + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = icl::mfma_16x16x4_f64(q[j], p[i], acc[i][j]);
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + (lane & 15);
                const int n = n0 + wn + 16 * j + 4 * r + (lane >> 4);
                if (m < a.M && n < a.N) {
                    const double b = (double)a.B0[(long)n * a.ldb0 + m];
                    a.R[(long)n * a.ldr + m] = (float)(a.alpha * b - acc[i][j][r]);
                }
            }
}

// ---------------------------------------------------------------------------
// Blocked solve with inverted 64 x 64 diagonal blocks. Step b of the canonical
// forward substitution, for every group of 64 right-hand sides (lines):
//   X_b = inv(L_bb) * B_b                          (64 x 64 x 64 FP32 MFMA)
//   B_r -= L_rb * X_b   for the trailing rows r    (128 rows per workgroup)
// Each workgroup of the step recomputes X_b for its lines from B_b, which no
// workgroup of the step writes, so the workgroups are independent; the one with
// group 0 also stores X_b to the dense solution X. L, B may be strided either
// way (canonical views of all side/uplo/trans variants).
// ---------------------------------------------------------------------------
constexpr int kStepCols = 64;  // lines per workgroup
constexpr int kStepRows = 128; // trailing rows per workgroup
// BX and Lr may be written in transposed order (k fastest across lanes): odd
// pitches keep those writes free of bank conflicts. D is written as float4.
constexpr int kLdD = kNB + 4, kLdB = kStepCols + 1, kLdL = kStepRows + 1;
constexpr size_t kStepLds = sizeof(float) * ((size_t)kNB * kLdD + (size_t)kNB * kLdB + (size_t)kNB * kLdL);

// A KN x R tile, element (r, k) = src[r * s_r + k * s_k] (zero beyond rmax /
// kmax), staged through registers: load() issues every read of the tile before
// any result is used, so the tile costs one memory latency, not one per
// element. RFAST: s_r is the unit stride (+1 or -1), so consecutive lanes take
// consecutive r; otherwise consecutive k. The reads are coalesced either way.
// Element i of a thread is (r_t, k_t + kStep * i) when RFAST and
// (r_t + kStep * i, k_t) otherwise, so its address is one per-thread base plus a
// workgroup-uniform multiple of the slow stride.
template <int KN, int R, bool RFAST>
struct TileRegs {
    static constexpr int kN = KN * R / 256;
    static constexpr int kStep = 256 / (RFAST ? R : KN);
    static_assert(256 % R == 0 && 256 % KN == 0, "tile edges must divide 256");
    float v[kN];

    __device__ __forceinline__ static void coords(int tid, int& r, int& k)
    {
        r = RFAST ? tid % R : tid / KN;
        k = RFAST ? tid / R : tid % KN;
    }
    // Elements outside the bounds read *safe (any valid address) and keep 0, so
    // the reads are straight-line code with no branch between them.
    __device__ __forceinline__ void load(const float* src, long s_r, long s_k, int rmax, int kmax, int tid,
                                         const float* safe)
    {
        int r, k;
        coords(tid, r, k);
        const float* p = src + (long)r * s_r + (long)k * s_k;
        const long slow_stride = (long)kStep * (RFAST ? s_k : s_r);
        const bool fast_ok = RFAST ? r < rmax : k < kmax;
        const int slow0 = RFAST ? k : r, slow_max = RFAST ? kmax : rmax;
#pragma unroll
        for (int i = 0; i < kN; ++i) {
            const bool ok = fast_ok && slow0 + kStep * i < slow_max;
            const float x = *(ok ? p + i * slow_stride : safe);
            v[i] = ok ? x : 0.0f;
        }
    }
    // T[k * LDT + r] = element (r, k)
    template <int LDT>
    __device__ __forceinline__ void store(float* T, int tid) const
    {
        int r, k;
        coords(tid, r, k);
#pragma unroll
        for (int i = 0; i < kN; ++i) T[(RFAST ? k + kStep * i : k) * LDT + (RFAST ? r : r + kStep * i)] = v[i];
    }
};

// Element r of accumulator j of the trailing update sits at (row_t + dr, col_t +
// dc) relative to (rb, c0), with (row_t, col_t) per thread and (dr, dc)
// compile-time constants: the lanes run along whichever of rows / lines is
// contiguous in B (32 x 32 MFMA output layout, rows 8 (r / 4) + 4 (lane / 32) + r % 4).
template <bool LINES_CONTIG>
__device__ __forceinline__ void update_base(int wave, int lane, int& row_t, int& col_t)
{
    row_t = 32 * wave + (LINES_CONTIG ? 4 * (lane >> 5) : (lane & 31));
    col_t = LINES_CONTIG ? (lane & 31) : 4 * (lane >> 5);
}
template <bool LINES_CONTIG>
__device__ __forceinline__ constexpr int update_dr(int j, int r)
{
    return LINES_CONTIG ? 8 * (r >> 2) + (r & 3) : 0;
}
template <bool LINES_CONTIG>
__device__ __forceinline__ constexpr int update_dc(int j, int r)
{
    return 32 * j + (LINES_CONTIG ? 0 : 8 * (r >> 2) + (r & 3));
}

// LINES_CONTIG: the lines of B are its contiguous dimension (right-side solves),
// so the trailing update maps lanes to lines for coalesced read-modify-writes.
// L_ROWS_CONTIG: the rows of the canonical L are its contiguous dimension
// (l_si is +1 or -1).
template <bool LINES_CONTIG, bool L_ROWS_CONTIG>
__global__ void __launch_bounds__(256) trsm_step_kernel(icl_trsm_blocked_args a, int blk, int groups)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    float* D = sm;                 // D[k][r]  = inv(L_bb)(r, k)
    float* BX = D + kNB * kLdD;    // BX[k][c] = B_b(k, c), then X_b(k, c)
    float* Lr = BX + kNB * kLdB;   // Lr[k][r] = L(rb + r, r0 + k)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ncb = (a.nlines + kStepCols - 1) / kStepCols;
    const int cb = blockIdx.x % ncb, g = blockIdx.x / ncb;
    const int r0 = blk * kNB, c0 = cb * kStepCols;
    const int rb = r0 + kNB + g * kStepRows;
    const int ncols = min(kStepCols, a.nlines - c0);
    const bool update = rb < a.nsys;

    // Every global read of the workgroup is issued up front: inv(L_bb) (dense
    // 64 x 64, 16-byte aligned), B_b, L_rb and the block of B this workgroup
    // updates, which no other workgroup of the step touches.
    // Out-of-range elements (all of L_rb and of the trailing block when there
    // are no trailing rows) read a valid element of B_b or L_bb instead.
    const float* b_org = a.B + (long)r0 * a.x_si + (long)c0 * a.x_sj;
    const float* l_diag = a.L + (long)r0 * (a.l_si + a.l_sk);
    TileRegs<kNB, kStepCols, LINES_CONTIG> bt;
    bt.load(b_org, a.x_sj, a.x_si, ncols, a.nsys - r0, tid, b_org);
    TileRegs<kNB, kStepRows, L_ROWS_CONTIG> lt;
    lt.load(a.L + (long)rb * a.l_si + (long)r0 * a.l_sk, a.l_si, a.l_sk, a.nsys - rb, a.nsys - r0, tid, l_diag);
    int row_t, col_t;
    update_base<LINES_CONTIG>(wave, lane, row_t, col_t);
    float* const pb = a.B + (long)(rb + row_t) * a.x_si + (long)(c0 + col_t) * a.x_sj;
    float old[2][16];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int dr = update_dr<LINES_CONTIG>(j, r), dc = update_dc<LINES_CONTIG>(j, r);
            const bool ok = rb + row_t + dr < a.nsys && col_t + dc < ncols;
            const float x = *(ok ? pb + dr * a.x_si + dc * a.x_sj : b_org);
            old[j][r] = ok ? x : 0.0f;
        }
    const float4* inv4 = reinterpret_cast<const float4*>(a.invD + (long)blk * kNB * kNB);
    float4 dv[kNB * kNB / 4 / 256];
#pragma unroll
    for (int i = 0; i < kNB * kNB / 4 / 256; ++i) dv[i] = inv4[tid + 256 * i];
    bt.template store<kLdB>(BX, tid);
    if (update) lt.template store<kLdL>(Lr, tid);
#pragma unroll
    for (int i = 0; i < kNB * kNB / 4 / 256; ++i) {
        const int e = 4 * (tid + 256 * i);
        *reinterpret_cast<float4*>(D + (e / kNB) * kLdD + (e % kNB)) = dv[i];
    }
    __syncthreads();

    // X_b quadrant of this wave: rows wr.., lines wc.. (lane <-> row)
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
    f32x16 x = f32x16{};
#pragma unroll 8
    for (int kk = 0; kk < kNB; kk += 2) {
        const int kr = kk + (lane >> 5);
        x = icl::mfma_32x32x2_f32(BX[kr * kLdB + wc + (lane & 31)], D[kr * kLdD + wr + (lane & 31)], x);
    }
    __syncthreads(); // every wave is done reading B_b
    {
        const int row = wr + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int col = wc + icl::mfma32_row(lane, r);
            BX[row * kLdB + col] = x[r];
            if (g == 0 && r0 + row < a.nsys && col < ncols) a.X[(long)(r0 + row) + (long)(c0 + col) * a.ldx] = x[r];
        }
    }
    if (!update) return;
    __syncthreads();

    // B(rb + row, c0 + col) -= sum_k L(rb + row, r0 + k) X_b(k, col); wave owns 32 rows, both line halves
    f32x16 u[2] = {f32x16{}, f32x16{}};
#pragma unroll 4
    for (int kk = 0; kk < kNB; kk += 2) {
        const int kr = kk + (lane >> 5);
        const float l = Lr[kr * kLdL + 32 * wave + (lane & 31)];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float xv = BX[kr * kLdB + 32 * j + (lane & 31)];
            u[j] = LINES_CONTIG ? icl::mfma_32x32x2_f32(l, xv, u[j]) : icl::mfma_32x32x2_f32(xv, l, u[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int dr = update_dr<LINES_CONTIG>(j, r), dc = update_dc<LINES_CONTIG>(j, r);
            if (rb + row_t + dr < a.nsys && col_t + dc < ncols) pb[dr * a.x_si + dc * a.x_sj] = old[j][r] - u[j][r];
        }
}

// B(i, j) = alpha * X(i, j) through the canonical view of B.
__global__ void trsm_scatter_kernel(icl_trsm_blocked_args a, float alpha)
{
    const long total = (long)a.nsys * a.nlines;
    const bool rows_fast = a.x_si == 1 || a.x_si == -1;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        long i, j;
        if (rows_fast) {
            j = e / a.nsys;
            i = e - j * a.nsys;
        } else {
            i = e / a.nlines;
            j = e - i * a.nlines;
        }
        a.B[i * a.x_si + j * a.x_sj] = alpha * a.X[i + j * a.ldx];
    }
}

// ---------------------------------------------------------------------------
// Elementwise 2-D updates.
// ---------------------------------------------------------------------------
__global__ void axpby_kernel(int m, int n, float alpha, const float* X, long ldx, float beta, float* Y, long ldy)
{
    const long total = (long)m * n;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long c = e / m, r = e - c * m;
        float v = (alpha != 0.0f) ? alpha * X[c * ldx + r] : 0.0f;
        if (beta != 0.0f) v = fmaf(beta, Y[c * ldy + r], v);
        Y[c * ldy + r] = v;
    }
}

__global__ void caxpby_kernel(int m, int n, float ar, float ai, const float* X, long ldx, float br, float bi,
                              float* Y, long ldy)
{
    const long total = (long)m * n;
    const bool a_zero = ar == 0.0f && ai == 0.0f;
    const bool b_zero = br == 0.0f && bi == 0.0f;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long c = e / m, r = e - c * m;
        float vr = 0.0f, vi = 0.0f;
        if (!a_zero) {
            const float xr = X[2 * (c * ldx + r)], xi = X[2 * (c * ldx + r) + 1];
            vr = fmaf(ar, xr, -ai * xi);
            vi = fmaf(ar, xi, ai * xr);
        }
        if (!b_zero) {
            const float yr = Y[2 * (c * ldy + r)], yi = Y[2 * (c * ldy + r) + 1];
            vr += fmaf(br, yr, -bi * yi);
            vi += fmaf(br, yi, bi * yr);
        }
        Y[2 * (c * ldy + r)] = vr;
        Y[2 * (c * ldy + r) + 1] = vi;
    }
}

unsigned elementwise_grid(long total)
{
    long blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

} // namespace

extern "C" hipError_t icl_trsm_diag(const icl_trsm_diag_args* a, hipStream_t stream)
{
    if (a->nb <= 0 || a->nlines <= 0) return hipSuccess;
    if (a->nb > kNB) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)((a->nlines + kNB - 1) / kNB);
    hipLaunchKernelGGL(trsm_diag_kernel, dim3(blocks), dim3(64), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t icl_trsm_invert_blocks(const icl_trsm_blocked_args* a, hipStream_t stream)
{
    if (a->nsys <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((a->nsys + kNB - 1) / kNB);
    hipLaunchKernelGGL(trsm_invert_kernel, dim3(blocks), dim3(64), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t icl_trsm_step(const icl_trsm_blocked_args* a, int block, hipStream_t stream)
{
    const int r0 = block * kNB;
    if (a->nlines <= 0 || r0 >= a->nsys) return hipSuccess;
    const int ncb = (a->nlines + kStepCols - 1) / kStepCols;
    const int below = a->nsys - r0 - kNB;
    const int groups = below > 0 ? (below + kStepRows - 1) / kStepRows : 1;
    using Kern = void (*)(icl_trsm_blocked_args, int, int);
    static const Kern kernels[4] = {trsm_step_kernel<false, false>, trsm_step_kernel<false, true>,
                                    trsm_step_kernel<true, false>, trsm_step_kernel<true, true>};
    static const hipError_t attr = [] {
        for (Kern k : kernels) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kStepLds);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }();
    if (attr != hipSuccess) return attr;
    const bool lines_contig = a->x_sj == 1 || a->x_sj == -1;
    const bool l_rows_contig = a->l_si == 1 || a->l_si == -1;
    const Kern kern = kernels[2 * lines_contig + l_rows_contig];
    hipLaunchKernelGGL(kern, dim3((unsigned)(ncb * groups)), dim3(256), kStepLds, stream, *a, block, groups);
    return hipGetLastError();
}

extern "C" hipError_t icl_trsm_scatter(const icl_trsm_blocked_args* a, float alpha, hipStream_t stream)
{
    if (a->nsys <= 0 || a->nlines <= 0) return hipSuccess;
    hipLaunchKernelGGL(trsm_scatter_kernel, dim3(elementwise_grid((long)a->nsys * a->nlines)), dim3(256), 0, stream,
                       *a, alpha);
    return hipGetLastError();
}

extern "C" hipError_t icl_residual_f64(const icl_residual_args* a, hipStream_t stream)
{
    if (a->M <= 0 || a->N <= 0) return hipSuccess;
    const long tiles = (long)((a->M + RT - 1) / RT) * ((a->N + RT - 1) / RT);
    hipLaunchKernelGGL(residual_f64_kernel, dim3((unsigned)tiles), dim3(256), 0, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t icl_axpby_2d(int m, int n, float alpha, const float* X, long ldx, float beta, float* Y,
                                   long ldy, hipStream_t stream)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(axpby_kernel, dim3(elementwise_grid((long)m * n)), dim3(256), 0, stream, m, n, alpha, X, ldx,
                       beta, Y, ldy);
    return hipGetLastError();
}

extern "C" hipError_t icl_caxpby_2d(int m, int n, float alpha_re, float alpha_im, const float* X, long ldx,
                                    float beta_re, float beta_im, float* Y, long ldy, hipStream_t stream)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(caxpby_kernel, dim3(elementwise_grid((long)m * n)), dim3(256), 0, stream, m, n, alpha_re,
                       alpha_im, X, ldx, beta_re, beta_im, Y, ldy);
    return hipGetLastError();
}
