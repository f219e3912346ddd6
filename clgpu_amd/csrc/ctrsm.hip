// CTRSM kernels: the solve of one diagonal block (at most 64 x 64), the
// inversion of every 64 x 64 diagonal block, the 256 x 256 panel inverses built
// from those, and the in-place product of a leaf with its block inverse.
//
// The substitution's arithmetic is the reference kernel's (libs/blas/funcs/Ctrsm/Ctrsm_naive.cl:16-183
// with the helpers of core/src/ocl/complex.h.cl:32-39): per pivot i,
// x_i = cdiv(b_i, a_ii) for a non-unit diagonal, then b_k -= cmul(a_ki, x_i) for
// every later k, and the solved value is stored as cmul(x_i, alpha). OP_C
// conjugates every element of A that is read, the diagonal included.
#include "mfma.hpp"
#include "kernel_abi.h"

namespace {

constexpr int kNB = 64; // largest diagonal block; one lane owns one right-hand side

__device__ __forceinline__ float2 c_mul(float2 a, float2 b)
{
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

__device__ __forceinline__ float2 c_div(float2 a, float2 b)
{
    const float d = fmaf(b.x, b.x, b.y * b.y);
    return make_float2(fmaf(a.x, b.x, a.y * b.y) / d, fmaf(a.y, b.x, -a.x * b.y) / d);
}

// The block staged in LDS transposed, Lt[i][k] = L(k, i) for k > i (conjugated
// for OP_C), so the update of pivot i reads one broadcast value per k; entries
// outside the block are 0 so the fully unrolled update never reads garbage.
struct CDiagBlock {
    float2 Lt[kNB][kNB + 1];
    float2 diag[kNB];
};

__global__ void __launch_bounds__(64) ctrsm_diag_kernel(icl_ctrsm_diag_args a)
{
    __shared__ CDiagBlock s;
    const int lane = threadIdx.x;
    const int nb = a.nb;
    const float2* L = reinterpret_cast<const float2*>(a.L);
    const float sg = a.conj ? -1.0f : 1.0f;
    for (int i = 0; i < kNB; ++i) {
        float2 v = make_float2(0.0f, 0.0f);
        if (i < nb && lane < nb && lane > i) {
            v = L[(long)lane * a.l_si + (long)i * a.l_sk];
            v.y *= sg;
        }
        s.Lt[i][lane] = v;
    }
    if (lane < nb) {
        float2 d = make_float2(1.0f, 0.0f);
        if (!a.unit) {
            d = L[(long)lane * (a.l_si + a.l_sk)];
            d.y *= sg;
        }
        s.diag[lane] = d;
    }
    __syncthreads();

    const long j = (long)blockIdx.x * kNB + lane;
    const bool active = j < a.nlines;
    float2* xcol = reinterpret_cast<float2*>(a.X) + j * a.x_sj;
    float2 x[kNB];
#pragma unroll
    for (int i = 0; i < kNB; ++i) x[i] = (active && i < nb) ? xcol[(long)i * a.x_si] : make_float2(0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < kNB; ++i) {
        if (i < nb) {
            float2 xi = x[i];
            if (!a.unit) xi = c_div(xi, s.diag[i]);
            x[i] = xi;
#pragma unroll
            for (int k = i + 1; k < kNB; ++k) {
                const float2 t = c_mul(s.Lt[i][k], xi);
                x[k].x -= t.x;
                x[k].y -= t.y;
            }
        }
    }
    if (active) {
        const float2 sp = make_float2(a.s_post_re, a.s_post_im);
#pragma unroll
        for (int i = 0; i < kNB; ++i)
            if (i < nb) xcol[(long)i * a.x_si] = c_mul(x[i], sp);
    }
}

} // namespace

extern "C" hipError_t icl_ctrsm_diag(const icl_ctrsm_diag_args* a, hipStream_t stream)
{
    if (a->nb <= 0 || a->nlines <= 0) return hipSuccess;
    if (a->nb > kNB) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)((a->nlines + kNB - 1) / kNB);
    hipLaunchKernelGGL(ctrsm_diag_kernel, dim3(blocks), dim3(64), 0, stream, *a);
    return hipGetLastError();
}

namespace {

// Inverts every 64 x 64 diagonal block of the canonical system (workgroup b
// handles unknowns [64 b, 64 b + nb)) and writes inv(L_bb) to
// invD[b * 4096 + j * 64 + i] (interleaved complex, column-major, rows and
// columns beyond nb zero). L(i, k) = L[i * l_si + k * l_sk], conjugated when
// conj. Recursive doubling: each wave inverts one 16 x 16 diagonal sub-block by
// substitution with the reference's complex multiply and divide (lane j forms
// column j, the sub-block read from LDS, no barrier per pivot), then the 32 x 32
// halves and the whole block follow from
// inv([A 0; C D]) = [inv(A) 0; -inv(D) C inv(A), inv(D)], each a pair of small
// complex products spread over the workgroup. Rows beyond nb are an identity
// block, so the padded inverse is exact there and those entries are written as
// zeros. The three LDS arrays (L, the inverse, the products' intermediate) take
// kCivLds bytes of dynamic LDS.
constexpr int kCivLd = kNB + 1;
constexpr size_t kCivLds = sizeof(float2) * (2 * kNB + kNB / 2) * kCivLd;
__global__ void __launch_bounds__(256) ctrsm_invert_kernel(icl_ctrsm_blocked_args a)
{
    extern __shared__ float sm[];
    float2 (*Ls)[kCivLd] = reinterpret_cast<float2 (*)[kCivLd]>(sm); // L(i, k), i >= k; identity beyond nb
    float2 (*Is)[kCivLd] = Ls + kNB;                                  // the inverse, built in place
    float2 (*Ts)[kCivLd] = Is + kNB;                                  // products' intermediate
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = blockIdx.x;
    const int r0 = b * kNB;
    const int nb = min(kNB, a.nsys - r0);
    const float2* L = reinterpret_cast<const float2*>(a.L) + (long)r0 * (a.l_si + a.l_sk);
    const float sg = a.conj ? -1.0f : 1.0f;
    {
        // 16 entries per thread, all loads in flight together: entry e = tid + 256 q
        // is (i, k) = (e / 64, e % 64); an entry outside the used triangle reads the
        // block's first diagonal element instead and is replaced. Each entry, the
        // diagonal included, has exactly one writer here.
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = tid + 256 * q, i = e / kNB, k = e % kNB;
            const bool used = i < nb && (k < i || (k == i && !a.unit));
            v[q] = L[used ? (long)i * a.l_si + (long)k * a.l_sk : 0];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = tid + 256 * q, i = e / kNB, k = e % kNB;
            const bool used = i < nb && (k < i || (k == i && !a.unit));
            Ls[i][k] = used ? make_float2(v[q].x, sg * v[q].y) : make_float2(k == i ? 1.0f : 0.0f, 0.0f);
            Is[i][k] = make_float2(0.0f, 0.0f);
        }
    }
    __syncthreads();
    // 16 x 16 diagonal sub-block w: lane j < 16 forms column j of its inverse.
    {
        const int o = 16 * wave, j = lane;
        if (j < 16) {
            float2 x[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                float2 acc = make_float2(i == j ? 1.0f : 0.0f, 0.0f);
#pragma unroll
                for (int k = 0; k < i; ++k) {
                    const float2 t = c_mul(Ls[o + i][o + k], x[k]);
                    acc.x -= t.x;
                    acc.y -= t.y;
                }
                x[i] = a.unit ? acc : c_div(acc, Ls[o + i][o + i]);
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) Is[o + i][o + j] = x[i];
        }
    }
    __syncthreads();
    // Halves: for the 32 x 32 blocks {0, 1} and {2, 3} (offsets 0 and 32), then the
    // whole block (offset 0, size 32): inv(D) C inv(A) with A at [o, o + h), D at
    // [o + h, o + 2 h), C = L[o + h .., o ..]; T = C inv(A), then -inv(D) T.
#pragma unroll
    for (int level = 0; level < 2; ++level) {
        const int h = level == 0 ? 16 : 32;          // size of A, C, D
        const int nblk = level == 0 ? 2 : 1;         // independent blocks
        const int per = (nblk * h * h + 255) / 256;  // elements per thread
#pragma unroll
        for (int step = 0; step < 2; ++step) {
            float2 out[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q >= per) break;
                const int e = tid + 256 * q;
                const int blk = e / (h * h), rc = e % (h * h), r = rc / h, c = rc % h;
                const int o = blk * 2 * h;
                float2 acc = make_float2(0.0f, 0.0f);
                if (step == 0) { // T(r, c) = sum_k C(r, k) inv(A)(k, c), k >= c
                    for (int k = c; k < h; ++k) {
                        const float2 t = c_mul(Ls[o + h + r][o + k], Is[o + k][o + c]);
                        acc.x += t.x;
                        acc.y += t.y;
                    }
                } else { // -sum_k inv(D)(r, k) T(k, c), k <= r
                    for (int k = 0; k <= r; ++k) {
                        const float2 t = c_mul(Is[o + h + r][o + h + k], Ts[blk * h + k][c]);
                        acc.x -= t.x;
                        acc.y -= t.y;
                    }
                }
                out[q] = acc;
            }
            __syncthreads(); // every read of Ts / Is of this step is done
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q >= per) break;
                const int e = tid + 256 * q;
                const int blk = e / (h * h), rc = e % (h * h), r = rc / h, c = rc % h;
                const int o = blk * 2 * h;
                if (step == 0) Ts[blk * h + r][c] = out[q];
                else Is[o + h + r][o + c] = out[q];
            }
            __syncthreads();
        }
    }
    float2* dst = reinterpret_cast<float2*>(a.invD) + (long)b * kNB * kNB;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int e = tid + 256 * q, j = e / kNB, i = e % kNB; // column j, row i: consecutive threads along i
        dst[e] = (i < nb && j < nb && i >= j) ? Is[i][j] : make_float2(0.0f, 0.0f);
    }
}

// X_b = s_post * inv(L_bb) B_b for one diagonal block and all lines, in place:
// B(i, j) = X[i * x_si + j * x_sj] (interleaved complex, i < nb). One workgroup
// of two waves per 32 lines (so that a 1024-line solve spreads over 32 CUs);
// wave w forms rows 32w..32w+31 of the product with v_mfma_f32_32x32x2_f32,
// Yr = Dr Br - Di Bi and Yi = Dr Bi + Di Br, the k range split in two halves
// with their own accumulators (four independent MFMA chains per wave).
// LINES_CONTIG: the lines are the contiguous dimension of B (right-side
// solves), so the result's lanes run along lines for coalesced stores.
constexpr int kApLines = 32;
constexpr int kApThreads = 128;
constexpr int kApLd = kNB + 1;
constexpr int kApLdB = kApLines + 1;
constexpr size_t kApLds = sizeof(float) * (2 * kNB * kApLd + 2 * kNB * kApLdB);
template <bool LINES_CONTIG>
__global__ void __launch_bounds__(kApThreads) ctrsm_apply_kernel(icl_ctrsm_apply_args a)
{
    extern __shared__ float sm[];
    float(*Dr)[kApLd] = reinterpret_cast<float(*)[kApLd]>(sm); // D[k][r] = inv(L_bb)(r, k)
    float(*Di)[kApLd] = Dr + kNB;
    float(*Br)[kApLdB] = reinterpret_cast<float(*)[kApLdB]>(sm + 2 * kNB * kApLd); // B[k][l] = B_b(k, l)
    float(*Bi)[kApLdB] = Br + kNB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = blockIdx.x * kApLines;
    const int ncols = min(kApLines, a.nlines - c0);
    const float2* inv = reinterpret_cast<const float2*>(a.invD);
    float2* X = reinterpret_cast<float2*>(a.X);
    // All global reads are issued before the first LDS store, so their latencies
    // overlap instead of adding up.
    constexpr int kInvPer = kNB * kNB / kApThreads, kBPer = kNB * kApLines / kApThreads;
    float2 iv[kInvPer], bv[kBPer];
#pragma unroll
    for (int i = 0; i < kInvPer; ++i) iv[i] = inv[tid + i * kApThreads]; // dense column-major: e = k * 64 + r
#pragma unroll
    for (int i = 0; i < kBPer; ++i) {
        const int e = tid + i * kApThreads;
        const int k = LINES_CONTIG ? e / kApLines : e % kNB, l = LINES_CONTIG ? e % kApLines : e / kNB;
        const bool in = k < a.nb && l < ncols;
        const long off = in ? (long)k * a.x_si + (long)(c0 + l) * a.x_sj : 0;
        bv[i] = in ? X[off] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int i = 0; i < kInvPer; ++i) {
        const int e = tid + i * kApThreads;
        Dr[e / kNB][e % kNB] = iv[i].x;
        Di[e / kNB][e % kNB] = iv[i].y;
    }
#pragma unroll
    for (int i = 0; i < kBPer; ++i) {
        const int e = tid + i * kApThreads;
        const int k = LINES_CONTIG ? e / kApLines : e % kNB, l = LINES_CONTIG ? e % kApLines : e / kNB;
        Br[k][l] = bv[i].x;
        Bi[k][l] = bv[i].y;
    }
    __syncthreads();
    const int wr = 32 * wave;
    f32x16 yr0 = f32x16{}, yi0 = f32x16{}, yr1 = f32x16{}, yi1 = f32x16{};
#pragma unroll 4
    for (int kk = 0; kk < kNB / 2; kk += 2) {
        const int k0 = kk + (lane >> 5), k1 = k0 + kNB / 2;
        const int rr = wr + (lane & 31), ll = lane & 31;
        const float dr0 = Dr[k0][rr], di0 = Di[k0][rr], br0 = Br[k0][ll], bi0 = Bi[k0][ll];
        const float dr1 = Dr[k1][rr], di1 = Di[k1][rr], br1 = Br[k1][ll], bi1 = Bi[k1][ll];
        if (LINES_CONTIG) { // lanes along lines
            yr0 = icl::mfma_32x32x2_f32(dr0, br0, yr0);
            yi0 = icl::mfma_32x32x2_f32(dr0, bi0, yi0);
            yr1 = icl::mfma_32x32x2_f32(dr1, br1, yr1);
            yi1 = icl::mfma_32x32x2_f32(dr1, bi1, yi1);
            yr0 = icl::mfma_32x32x2_f32(-di0, bi0, yr0);
            yi0 = icl::mfma_32x32x2_f32(di0, br0, yi0);
            yr1 = icl::mfma_32x32x2_f32(-di1, bi1, yr1);
            yi1 = icl::mfma_32x32x2_f32(di1, br1, yi1);
        } else { // lanes along rows
            yr0 = icl::mfma_32x32x2_f32(br0, dr0, yr0);
            yi0 = icl::mfma_32x32x2_f32(bi0, dr0, yi0);
            yr1 = icl::mfma_32x32x2_f32(br1, dr1, yr1);
            yi1 = icl::mfma_32x32x2_f32(bi1, dr1, yi1);
            yr0 = icl::mfma_32x32x2_f32(bi0, -di0, yr0);
            yi0 = icl::mfma_32x32x2_f32(br0, di0, yi0);
            yr1 = icl::mfma_32x32x2_f32(bi1, -di1, yr1);
            yi1 = icl::mfma_32x32x2_f32(br1, di1, yi1);
        }
    }
    const float2 sp = make_float2(a.s_post_re, a.s_post_im);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = wr + (LINES_CONTIG ? icl::mfma32_row(lane, r) : (lane & 31));
        const int col = LINES_CONTIG ? (lane & 31) : icl::mfma32_row(lane, r);
        if (row < a.nb && col < ncols)
            X[(long)row * a.x_si + (long)(c0 + col) * a.x_sj] =
                c_mul(make_float2(yr0[r] + yr1[r], yi0[r] + yi1[r]), sp);
    }
}

// Inverses of the 256 x 256 diagonal panels (see icl_ctrsm_invert_panels), by
// recursive doubling over 64 x 64 blocks B_ij of a panel (0 <= i, j < 4), from
// the inverted diagonal blocks iD_i: a lower triangular [A 0; C D] has the
// inverse [inv(A) 0; -inv(D) C inv(A) inv(D)]. Three launches:
//  level 0, workgroup (p, h): the 128 x 128 diagonal halves {2h, 2h + 1}:
//    W_aa = iD_a, W_dd = iD_d, W_ad = 0, W_da = -iD_d (L_da iD_a);
//    workgroup h = 0 also zeroes the upper right quadrant (W_02, W_03, W_12, W_13);
//  level 1, workgroup (p, q): T_rc = sum_{k = c..1} L_rk W_kc (r = 2 + q / 2, c = q % 2),
//    i.e. T = C inv(A), into a scratch array;
//  level 2, workgroup (p, q): W_rc = -sum_{k = 2..r} W_rk T_kc, i.e. -inv(D) T.
// The longest chain is two 64 x 64 x 64 complex products per launch, where
// building the column blocks one after the other (M_ij = -iD_i sum_{k=j}^{i-1}
// L_ik M_kj) chains nine. Each product runs on v_mfma_f32_32x32x2_f32 with the
// operands staged in LDS as planar re / im tiles, A as [k][row] and B as
// [k][col], each of the four waves owning a 32 x 32 quadrant.
constexpr int kPanRows = 256;
constexpr int kPnLd = kNB + 1;
constexpr size_t kPanInvLds = sizeof(float) * 4 * kNB * kPnLd;
constexpr int kPer = kNB * kNB / 256; // elements of a 64 x 64 block per thread; e = tid + 256 q: (x = e % 64, y = e / 64)

// A 64 x 64 operand block: element (x, y) of the block.
struct PanBlk {
    enum Kind { L, INVD, WBLK, TBLK } kind;
    int bi, bj; // block coordinates within the panel (INVD: diagonal block bi; TBLK: T_(bi, bj))
};

struct PanCtx {
    const icl_ctrsm_panel_inv_args& a;
    int pnl, ps, plen;
    float2* W;
    float2* T;
    __device__ long w_at(int r, int c) const
    {
        return a.rev ? (long)(plen - 1 - r) * a.w_rs + (long)(plen - 1 - c) * a.w_cs
                     : (long)r * a.w_rs + (long)c * a.w_cs;
    }
    __device__ float2* t_blk(int r, int c) const { return T + ((long)pnl * 4 + (r - 2) * 2 + c) * kNB * kNB; }
    // All loads first, unconditionally (an entry outside the system or the panel
    // reads element 0 of its array instead and is replaced by 0 afterwards).
    __device__ void load(const PanBlk& b, int tid, float2 (&v)[kPer]) const
    {
        const float2* L = reinterpret_cast<const float2*>(a.L);
        const float2* invD = reinterpret_cast<const float2*>(a.invD);
        const float sg = a.conj ? -1.0f : 1.0f;
        bool ok[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + 256 * q, x = e % kNB, y = e / kNB;
            const int r = kNB * b.bi + x, c = kNB * b.bj + y;
            switch (b.kind) {
            case PanBlk::L:
                ok[q] = ps + r < a.nsys && ps + c < a.nsys;
                v[q] = L[ok[q] ? (long)(ps + r) * a.l_si + (long)(ps + c) * a.l_sk : 0];
                v[q].y *= sg;
                break;
            case PanBlk::INVD:
                ok[q] = true;
                v[q] = invD[((long)(ps / kNB + b.bi) * kNB + y) * kNB + x];
                break;
            case PanBlk::WBLK:
                ok[q] = r < plen && c < plen;
                v[q] = W[ok[q] ? w_at(r, c) : 0];
                break;
            default:
                ok[q] = true;
                v[q] = t_blk(b.bi, b.bj)[(long)y * kNB + x];
                break;
            }
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q)
            if (!ok[q]) v[q] = make_float2(0.0f, 0.0f);
    }
};

__global__ void __launch_bounds__(256) ctrsm_invert_panels_kernel(icl_ctrsm_panel_inv_args a, int level)
{
    extern __shared__ float sm[];
    float(*Ar)[kPnLd] = reinterpret_cast<float(*)[kPnLd]>(sm); // A[k][row]
    float(*Ai)[kPnLd] = Ar + kNB;
    float(*Br)[kPnLd] = Ai + kNB; // B[k][col]
    float(*Bi)[kPnLd] = Br + kNB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1);
    const int pnl = blockIdx.x, ps = pnl * kPanRows;
    const int plen = min(kPanRows, a.nsys - ps), nb = (plen + kNB - 1) / kNB;
    const PanCtx cx{a, pnl, ps, plen, reinterpret_cast<float2*>(a.W) + (long)pnl * kPanRows * kPanRows,
                    reinterpret_cast<float2*>(a.T)};

    // Stage X as the A tile and Y as the B tile (after a barrier, so no wave still
    // reads the previous tiles), then acc (+)= X Y; lanes along rows of the result.
    auto product = [&](const PanBlk& X, const PanBlk& Y, f32x16& accr, f32x16& acci) {
        float2 va[kPer], vb[kPer];
        cx.load(X, tid, va);
        cx.load(Y, tid, vb);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + 256 * q, x = e % kNB, y = e / kNB; // X: (row x, k y); Y: (k x, col y)
            Ar[y][x] = va[q].x;
            Ai[y][x] = va[q].y;
            Br[x][y] = vb[q].x;
            Bi[x][y] = vb[q].y;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < kNB; kk += 2) {
            const int kr = kk + (lane >> 5);
            const float ar = Ar[kr][wr + (lane & 31)], ai = Ai[kr][wr + (lane & 31)];
            const float br = Br[kr][wc + (lane & 31)], bi = Bi[kr][wc + (lane & 31)];
            accr = icl::mfma_32x32x2_f32(br, ar, accr);
            acci = icl::mfma_32x32x2_f32(bi, ar, acci);
            accr = icl::mfma_32x32x2_f32(bi, -ai, accr);
            acci = icl::mfma_32x32x2_f32(br, ai, acci);
        }
    };
    // Element (row, col) of this lane's accumulator entry r.
    auto row_of = [&](int) { return wr + (lane & 31); };
    auto col_of = [&](int r) { return wc + icl::mfma32_row(lane, r); };
    auto store_w = [&](int bi, int bj, const f32x16& accr, const f32x16& acci, float s) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = kNB * bi + row_of(r), col = kNB * bj + col_of(r);
            if (row < plen && col < plen) cx.W[cx.w_at(row, col)] = make_float2(s * accr[r], s * acci[r]);
        }
    };
    auto copy_w = [&](const PanBlk& src, int bi, int bj) { // W_(bi, bj) = src (zero when src is null)
        float2 v[kPer];
        if (src.kind == PanBlk::INVD) cx.load(src, tid, v);
        else
            for (int q = 0; q < kPer; ++q) v[q] = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + 256 * q, r = kNB * bi + e % kNB, c = kNB * bj + e / kNB;
            if (r < plen && c < plen) cx.W[cx.w_at(r, c)] = v[q];
        }
    };
    const PanBlk zero{PanBlk::WBLK, -1, -1};

    if (level == 0) {
        const int h = blockIdx.y, da = 2 * h, dd = da + 1;
        if (da >= nb) return;
        copy_w({PanBlk::INVD, da, da}, da, da);
        copy_w(zero, da, dd);
        if (h == 0)
            for (int i = 0; i < 2; ++i)
                for (int j = 2; j < 4; ++j) copy_w(zero, i, j);
        if (dd >= nb) return;
        copy_w({PanBlk::INVD, dd, dd}, dd, dd);
        f32x16 sr = f32x16{}, si = f32x16{};
        product({PanBlk::L, dd, da}, {PanBlk::INVD, da, da}, sr, si); // S = L_da iD_a
        // W_da = -iD_d S: S becomes the B tile, iD_d the A tile.
        float2 vd[kPer];
        cx.load({PanBlk::INVD, dd, dd}, tid, vd);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            Br[row_of(r)][col_of(r)] = sr[r];
            Bi[row_of(r)][col_of(r)] = si[r];
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + 256 * q, x = e % kNB, y = e / kNB;
            Ar[y][x] = vd[q].x;
            Ai[y][x] = vd[q].y;
        }
        __syncthreads();
        f32x16 mr = f32x16{}, mi = f32x16{};
#pragma unroll 8
        for (int kk = 0; kk < kNB; kk += 2) {
            const int kr = kk + (lane >> 5);
            const float ar = Ar[kr][wr + (lane & 31)], ai = Ai[kr][wr + (lane & 31)];
            const float br = Br[kr][wc + (lane & 31)], bi = Bi[kr][wc + (lane & 31)];
            mr = icl::mfma_32x32x2_f32(br, ar, mr);
            mi = icl::mfma_32x32x2_f32(bi, ar, mi);
            mr = icl::mfma_32x32x2_f32(bi, -ai, mr);
            mi = icl::mfma_32x32x2_f32(br, ai, mi);
        }
        store_w(dd, da, mr, mi, -1.0f);
        return;
    }
    const int q = blockIdx.y, r = 2 + (q >> 1), c = q & 1;
    if (r >= nb) return;
    f32x16 accr = f32x16{}, acci = f32x16{};
    if (level == 1) { // T_rc = sum_{k = c..1} L_rk W_kc
        for (int k = c; k < 2; ++k) product({PanBlk::L, r, k}, {PanBlk::WBLK, k, c}, accr, acci);
        float2* t = cx.t_blk(r, c);
#pragma unroll
        for (int e = 0; e < 16; ++e) t[(long)col_of(e) * kNB + row_of(e)] = make_float2(accr[e], acci[e]);
        return;
    }
    // level 2: W_rc = -sum_{k = 2..r} W_rk T_kc
    for (int k = 2; k <= r; ++k) product({PanBlk::WBLK, r, k}, {PanBlk::TBLK, k, c}, accr, acci);
    store_w(r, c, accr, acci, -1.0f);
}

} // namespace

extern "C" hipError_t icl_ctrsm_invert_panels(const icl_ctrsm_panel_inv_args* a, hipStream_t stream)
{
    if (a->nsys <= 0) return hipSuccess;
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(ctrsm_invert_panels_kernel),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPanInvLds);
    if (attr != hipSuccess) return attr;
    const unsigned panels = (unsigned)((a->nsys + kPanRows - 1) / kPanRows);
    for (int level = 0; level < 3; ++level) {
        hipLaunchKernelGGL(ctrsm_invert_panels_kernel, dim3(panels, level == 0 ? 2 : 4), dim3(256), kPanInvLds, stream,
                           *a, level);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

extern "C" hipError_t icl_ctrsm_invert_blocks(const icl_ctrsm_blocked_args* a, hipStream_t stream)
{
    if (a->nsys <= 0) return hipSuccess;
    static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(ctrsm_invert_kernel),
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCivLds);
    if (attr != hipSuccess) return attr;
    const unsigned blocks = (unsigned)((a->nsys + kNB - 1) / kNB);
    hipLaunchKernelGGL(ctrsm_invert_kernel, dim3(blocks), dim3(256), kCivLds, stream, *a);
    return hipGetLastError();
}

extern "C" hipError_t icl_ctrsm_apply(const icl_ctrsm_apply_args* a, hipStream_t stream)
{
    if (a->nb <= 0 || a->nlines <= 0) return hipSuccess;
    if (a->nb > kNB) return hipErrorInvalidValue;
    static const hipError_t attr = [] {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(ctrsm_apply_kernel<true>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)kApLds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(ctrsm_apply_kernel<false>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kApLds);
        return e;
    }();
    if (attr != hipSuccess) return attr;
    const unsigned blocks = (unsigned)((a->nlines + kApLines - 1) / kApLines);
    if (a->x_sj == 1 || a->x_sj == -1)
        hipLaunchKernelGGL(ctrsm_apply_kernel<true>, dim3(blocks), dim3(kApThreads), kApLds, stream, *a);
    else
        hipLaunchKernelGGL(ctrsm_apply_kernel<false>, dim3(blocks), dim3(kApThreads), kApLds, stream, *a);
    return hipGetLastError();
}
