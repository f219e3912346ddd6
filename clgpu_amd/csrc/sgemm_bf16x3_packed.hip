// Split-BF16 (x3) SGEMM on pre-split, pre-packed operands.
//
// A prepass (split_pack_kernel, one launch for both operands) reads op(A)
// (m x k) and op(B)^T (n x k) once, splits every FP32 x into hi = bf16(x),
// lo = bf16(x - hi), and writes them in a slab-major layout that is exactly the
// LDS image the GEMM wants:
//
//   P[ks][r][h][kk]   ks = k / 32, r = row (m for A, n for B), h = 0 hi / 1 lo,
//                     kk = k % 32;  offset ((ks * Rpad + r) * 2 + h) * 32 + kk
//
// Rows are padded to the GEMM tile and K to 32 with zeros, so the GEMM main loop
// has no bounds checks, no transposes and no conversions: each 32-deep slab of a
// BM-row tile is one contiguous BM * 128-byte block, copied by the LDS DMA path
// into LDS rows of 64 bf16, from which every lane reads its 8-k fragments with
// 16-byte LDS loads for v_mfma_f32_32x32x16_bf16.
// acc += Bh*Ah + Bl*Ah + Bh*Al in FP32, as in the unpacked kernel.
#include "gemm_variants.hpp"

namespace icl {

constexpr int kPackK = 32; // k per slab

// ---------------------------------------------------------------------------
// Prepass: both operands in one launch. Workgroup (x, y) packs 64 rows of
// op(A) (x < ablocks) or of op(B)^T (the rest) over k = 64 y .. 64 y + 63, i.e.
// two slabs, staged through LDS. Element (r, k) of a source is X[k + r * ld]
// when kcontig, else X[r + k * ld]; the branch on it is uniform per workgroup.
// All 16 reads of a thread are issued before any is stored.
// ---------------------------------------------------------------------------
struct PackSrc {
    const float* X;
    long ld;
    int R, K, Rpad, kcontig; // Rpad: rows of the packed layout (its slab stride)
    __bf16* P;
    int blocks;              // 64-row blocks packed from X / P (rows R.. of them are zero padding)
};

__global__ void __launch_bounds__(256) split_pack_kernel(PackSrc a, PackSrc b, int ablocks, int slabs)
{
    __shared__ float T[64][64 + 1];
    const bool is_a = (int)blockIdx.x < ablocks;
    const PackSrc o = is_a ? a : b;
    const int r0 = (is_a ? (int)blockIdx.x : (int)blockIdx.x - ablocks) * 64;
    const int k0 = blockIdx.y * 64;
    const int t = threadIdx.x;
    float v[16];
    if (o.kcontig) {
        const int k = t & 63, rb = t >> 6;
        const bool kok = k0 + k < o.K;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = rb + 4 * i;
            v[i] = (kok && r0 + r < o.R) ? o.X[(long)(r0 + r) * o.ld + k0 + k] : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) T[rb + 4 * i][k] = v[i];
    } else {
        const int r = t & 63, kb = t >> 6;
        const bool rok = r0 + r < o.R;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int k = kb + 4 * i;
            v[i] = (rok && k0 + k < o.K) ? o.X[(long)(k0 + k) * o.ld + r0 + r] : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) T[r][kb + 4 * i] = v[i];
    }
    __syncthreads();
    // thread -> row r, 16 consecutive k from kq: slab 2y + kq / 32, in-slab offset kq % 32
    const int r = t >> 2, kq = (t & 3) * 16;
    const int ks = 2 * (int)blockIdx.y + kq / kPackK;
    if (ks >= slabs) return;
    bf16x8 hi[2], lo[2];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        __bf16 h, l;
        split_bf16(T[r][kq + i], h, l);
        hi[i / 8][i % 8] = h;
        lo[i / 8][i % 8] = l;
    }
    __bf16* dst = o.P + ((size_t)ks * o.Rpad + r0 + r) * (2 * kPackK) + kq % kPackK;
    *reinterpret_cast<bf16x8*>(dst) = hi[0];
    *reinterpret_cast<bf16x8*>(dst + 8) = hi[1];
    *reinterpret_cast<bf16x8*>(dst + kPackK) = lo[0];
    *reinterpret_cast<bf16x8*>(dst + kPackK + 8) = lo[1];
}

struct PackedArgs {
    icl_gemm_args g; // M, N, C, ldc, alpha, beta, mask (A/B/K fields unused here)
    const __bf16* Ap; // row 0 of this launch in the packed op(A)
    const __bf16* Bp;
    int Mpad, Npad, slabs;  // tile grid of this launch (M, N padded), slabs of K
    int a_rows, b_rows;     // rows of the packed layouts (their slab strides)
};

// ---------------------------------------------------------------------------
// GEMM on the packed operands. Slab copies use the LDS DMA path
// (global_load_lds_dwordx4): the next slab streams straight into the other LDS
// stage while the current one is multiplied, without passing through VGPRs.
// One DMA instruction writes 64 lanes x 16 bytes = 1 KB contiguous LDS (8 rows of
// 128 bytes), so rows are unpadded and bank conflicts on the 16-byte fragment
// reads are avoided by storing logical 16-byte chunk c of row r at physical
// chunk c ^ lds_swizzle(r) (the DMA lanes gather the logical chunks from global
// memory accordingly).
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(1))) void global_void;
typedef __attribute__((address_space(3))) void lds_void;

template <int BM, int BN, int WVM, int WVN>
struct DmaCfg {
    static constexpr int NW = WVM * WVN;
    static constexpr int NT = NW * 64;
    static constexpr int WM = BM / WVM, WN = BN / WVN;
    static constexpr int FM = WM / 32, FN = WN / 32;
    static constexpr int ROWB = 2 * kPackK * 2;            // 128 bytes per LDS row
    static constexpr int STAGE_BYTES = (BM + BN) * ROWB;
    static constexpr int IA = BM / 8 / NW, IB = BN / 8 / NW; // DMA instructions per wave per slab
    static constexpr size_t LDS_BYTES = 2 * (size_t)STAGE_BYTES;
    static_assert((BM / 8) % NW == 0 && (BN / 8) % NW == 0, "rows must divide among waves");
    static_assert(WM % 32 == 0 && WN % 32 == 0, "bad wave tile");
};

// L3: the level-3 instantiation -- tiles of the written triangle only
// (write_tri), the slabs of the tile's k window only (kband), and the epilogue
// that stores only the triangle; the plain instantiation is compiled without it.
template <int BM, int BN, int WVM, int WVN, bool EDGE, bool L3 = false>
__global__ void __launch_bounds__(WVM* WVN * 64) sgemm_bf16x3_dma_kernel(PackedArgs p)
{
    using Cfg = DmaCfg<BM, BN, WVM, WVN>;
    extern __shared__ __attribute__((aligned(1024))) unsigned char ldsd[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm0 = (wave % WVM) * Cfg::WM;
    const int wn0 = (wave / WVM) * Cfg::WN;
    int tm, tn;
    if constexpr (L3) {
        if (!tile_coords_level3<BM, BN>(p.g.write_tri, p.g.kband, p.Mpad / BM, p.Npad / BN, tm, tn)) return;
    } else
        tile_coords<8>(p.Mpad / BM, p.Npad / BN, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    int ks0 = 0, ks1 = p.slabs; // slabs [ks0, ks1) of the tile's k window
    if constexpr (L3) {
        int kb = 0, ke = 0;
        if (!tile_window<BM, BN>(p.g, m0, n0, kb, ke)) return;
        ks0 = kb / kPackK;
        ks1 = ke > kb ? (ke + kPackK - 1) / kPackK : ks0;
    }

    const size_t a_slab = (size_t)p.a_rows * (2 * kPackK);
    const size_t b_slab = (size_t)p.b_rows * (2 * kPackK);
    const __bf16* a_src = p.Ap + (size_t)m0 * (2 * kPackK);
    const __bf16* b_src = p.Bp + (size_t)n0 * (2 * kPackK);
    // lane -> (row within an 8-row DMA block, logical chunk)
    const int lrow = lane >> 3;

    auto issue = [&](int ks, int s) {
        unsigned char* stage = ldsd + s * Cfg::STAGE_BYTES;
        const __bf16* a = a_src + ks * a_slab;
        const __bf16* b = b_src + ks * b_slab;
#pragma unroll
        for (int i = 0; i < Cfg::IA; ++i) {
            const int r0 = 8 * (wave + Cfg::NW * i);
            const __bf16* g = a + (size_t)(r0 + lrow) * (2 * kPackK) + ((lane & 7) ^ lds_swizzle(r0 + lrow)) * 8;
            __builtin_amdgcn_global_load_lds((global_void*)g, (lds_void*)(stage + r0 * Cfg::ROWB), 16, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < Cfg::IB; ++i) {
            const int r0 = 8 * (wave + Cfg::NW * i);
            const __bf16* g = b + (size_t)(r0 + lrow) * (2 * kPackK) + ((lane & 7) ^ lds_swizzle(r0 + lrow)) * 8;
            __builtin_amdgcn_global_load_lds((global_void*)g, (lds_void*)(stage + (BM + r0) * Cfg::ROWB), 16, 0, 0);
        }
    };

    f32x16 acc[Cfg::FM][Cfg::FN];
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = f32x16{};

    if (ks0 < ks1) {
        issue(ks0, 0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int ks = ks0; ks < ks1; ++ks) {
            const int cur = (ks - ks0) & 1;
            if (ks + 1 < ks1) issue(ks + 1, cur ^ 1);
            const unsigned char* As = ldsd + cur * Cfg::STAGE_BYTES;
            const unsigned char* Bs = As + BM * Cfg::ROWB;
#pragma unroll
            for (int kk = 0; kk < kPackK; kk += 16) {
                const int c = (kk >> 3) + (lane >> 5); // logical hi chunk 0..3; lo = c + 4
                bf16x8 ah[Cfg::FM], al[Cfg::FM], bh[Cfg::FN], bl[Cfg::FN];
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i) {
                    const int r = wm0 + 32 * i + (lane & 31);
                    const unsigned char* row = As + r * Cfg::ROWB;
                    ah[i] = *reinterpret_cast<const bf16x8*>(row + ((c ^ lds_swizzle(r)) << 4));
                    al[i] = *reinterpret_cast<const bf16x8*>(row + (((c + 4) ^ lds_swizzle(r)) << 4));
                }
#pragma unroll
                for (int j = 0; j < Cfg::FN; ++j) {
                    const int r = wn0 + 32 * j + (lane & 31);
                    const unsigned char* row = Bs + r * Cfg::ROWB;
                    bh[j] = *reinterpret_cast<const bf16x8*>(row + ((c ^ lds_swizzle(r)) << 4));
                    bl[j] = *reinterpret_cast<const bf16x8*>(row + (((c + 4) ^ lds_swizzle(r)) << 4));
                }
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = mfma_32x32x16_bf16(bh[j], ah[i], acc[i][j]);
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = mfma_32x32x16_bf16(bl[j], ah[i], acc[i][j]);
#pragma unroll
                for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
                    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = mfma_32x32x16_bf16(bh[j], al[i], acc[i][j]);
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
        for (int j = 0; j < Cfg::FN; ++j)
            store_real_tile<EDGE, L3>(p.g, acc[i][j], m0 + wm0 + 32 * i + (lane & 31), n0 + wn0 + 32 * j, lane);
}

namespace {

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

template <int BM, int BN>
size_t packed_workspace(const icl_gemm_args* g)
{
    const size_t kpad = (size_t)round_up(g->K, kPackK);
    return ((size_t)round_up(g->M, BM) + (size_t)round_up(g->N, BN)) * kpad * 2 * sizeof(__bf16);
}

hipError_t pack(const PackSrc& a, const PackSrc& b, int slabs, hipStream_t s)
{
    const dim3 grid((unsigned)(a.blocks + b.blocks), (unsigned)((slabs + 1) / 2));
    hipLaunchKernelGGL(split_pack_kernel, grid, dim3(256), 0, s, a, b, a.blocks, slabs);
    return hipGetLastError();
}

// The rows [r0, r0 + rows) of a source and its packed layout as a PackSrc of
// their own (r0 a multiple of 64; rows up to rpad_rows are padding).
PackSrc pack_rows(const PackSrc& whole, int r0, int rows, int rpad_rows)
{
    PackSrc s = whole;
    s.X = whole.X + (whole.kcontig ? (long)r0 * whole.ld : (long)r0);
    s.R = rows;
    s.P = whole.P + (size_t)r0 * (2 * kPackK);
    s.blocks = rpad_rows / 64;
    return s;
}

template <int BM, int BN, int WVM, int WVN>
hipError_t packed_launch(const icl_gemm_args* g, int opa, int opb, int edge, void* ws, hipStream_t s)
{
    if (!ws) return hipErrorInvalidValue;
    PackedArgs p{};
    p.g = *g;
    p.Mpad = round_up(g->M, BM);
    p.Npad = round_up(g->N, BN);
    p.slabs = round_up(g->K, kPackK) / kPackK;
    p.a_rows = p.Mpad;
    p.b_rows = p.Npad;
    __bf16* Ap = static_cast<__bf16*>(ws);
    __bf16* Bp = Ap + (size_t)p.Mpad * p.slabs * (2 * kPackK);
    p.Ap = Ap;
    p.Bp = Bp;
    // op(A)(m, k): N -> A[m + k*lda] (rows contiguous), T/C -> A[k + m*lda] (k contiguous)
    const PackSrc pa{g->A, g->lda, g->M, g->K, p.Mpad, opa != 0, Ap, p.Mpad / 64};
    // op(B)(k, n) as rows n: N -> B[k + n*ldb] (k contiguous), T/C -> B[n + k*ldb]
    const PackSrc pb{g->B, g->ldb, g->N, g->K, p.Npad, opb == 0, Bp, p.Npad / 64};
    using C = DmaCfg<BM, BN, WVM, WVN>;
    (void)edge;
    hipError_t e;

    // Large plain products with a side stream: the operands in halves (rows of
    // op(A), rows of op(B)^T). The handle's stream packs A0 and B0 and multiplies
    // A0 x B0 while the side stream packs A1 and B1; then A0 x B1 and A1 x B
    // follow, once the side stream is done. Only the first halves' pack is
    // exposed, instead of the whole prepass. Each launch computes its elements
    // exactly as the single launch would (same slabs, same order).
    const icl_side_lane* side = icl_gemm_side_lane();
    const int m0 = round_up(g->M / 2, BM), n0 = round_up(g->N / 2, BN);
    if (side && !uses_level3(g) && g->M % BM == 0 && g->N % BN == 0 && m0 < g->M && n0 < g->N) {
        // The side stream starts after the first halves' pack, so the two packs do
        // not share the memory bandwidth: the second runs beside A0 x B0, which
        // leaves registers and LDS for its workgroups on every CU.
        if ((e = pack(pack_rows(pa, 0, m0, m0), pack_rows(pb, 0, n0, n0), p.slabs, s)) != hipSuccess) return e;
        if ((e = hipEventRecord(side->start, s)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(side->stream, side->start, 0)) != hipSuccess) return e;
        if ((e = pack(pack_rows(pa, m0, g->M - m0, p.Mpad - m0), pack_rows(pb, n0, g->N - n0, p.Npad - n0), p.slabs,
                      side->stream)) != hipSuccess)
            return e;
        if ((e = hipEventRecord(side->done, side->stream)) != hipSuccess) return e;
        if ((e = icl_timer_mark(s)) != hipSuccess) return e;
        // (row offset, rows, column offset, columns) of C per launch
        const int part[3][4] = {{0, m0, 0, n0}, {0, m0, n0, g->N - n0}, {m0, g->M - m0, 0, g->N}};
        for (int q = 0; q < 3; ++q) {
            if (q == 1 && (e = hipStreamWaitEvent(s, side->done, 0)) != hipSuccess) return e;
            icl_gemm_args gq = *g;
            gq.M = part[q][1];
            gq.N = part[q][3];
            gq.C = g->C + (long)part[q][2] * g->ldc + part[q][0];
            PackedArgs pq = p;
            pq.g = gq;
            pq.Mpad = gq.M;
            pq.Npad = gq.N;
            pq.Ap = Ap + (size_t)part[q][0] * (2 * kPackK);
            pq.Bp = Bp + (size_t)part[q][2] * (2 * kPackK);
            e = launch_gemm<BM, BN, C::NT, C::LDS_BYTES>(sgemm_bf16x3_dma_kernel<BM, BN, WVM, WVN, false>, &gq, s, pq);
            if (e != hipSuccess) return e;
        }
        return icl_timer_mark(s);
    }

    e = pack(pa, pb, p.slabs, s);
    if (e != hipSuccess) return e;
    const bool even = g->M % BM == 0 && g->N % BN == 0;
    // launch_gemm sizes the grid from (M, N); the packed grid is the same.
    if ((e = icl_timer_mark(s)) != hipSuccess) return e;
    if (uses_level3(g))
        e = even ? launch_gemm<BM, BN, C::NT, C::LDS_BYTES>(sgemm_bf16x3_dma_kernel<BM, BN, WVM, WVN, false, true>, g,
                                                           s, p)
                 : launch_gemm<BM, BN, C::NT, C::LDS_BYTES>(sgemm_bf16x3_dma_kernel<BM, BN, WVM, WVN, true, true>, g,
                                                           s, p);
    else if (even)
        e = launch_gemm<BM, BN, C::NT, C::LDS_BYTES>(sgemm_bf16x3_dma_kernel<BM, BN, WVM, WVN, false>, g, s, p);
    else
        e = launch_gemm<BM, BN, C::NT, C::LDS_BYTES>(sgemm_bf16x3_dma_kernel<BM, BN, WVM, WVN, true>, g, s, p);
    if (e != hipSuccess) return e;
    return icl_timer_mark(s);
}

} // namespace
} // namespace icl

// TF: measured useful TFLOP/s at m = n = k = 8192 (NN), pack prepass included; OCC: workgroups per CU.
#define ICL_SGEMM_BF16X3_PACKED(name, BM, BN, WVM, WVN, TF, OCC)                                                       \
    {name,                                                                                                             \
     ICL_KIND_SGEMM,                                                                                                   \
     ICL_MATH_BF16X3,                                                                                                  \
     BM,                                                                                                               \
     BN,                                                                                                               \
     icl::kPackK,                                                                                                      \
     64 * WVM * WVN,                                                                                                   \
     icl::packed_launch<BM, BN, WVM, WVN>,                                                                             \
     icl::packed_workspace<BM, BN>,                                                                                    \
     TF,                                                                                                               \
     OCC,                                                                                                              \
     1}

static const icl_gemm_variant kTable[] = {
    ICL_SGEMM_BF16X3_PACKED("sgemm_bf16x3d_mfma_256x256w8", 256, 256, 4, 2, 458.0f, 1),
    ICL_SGEMM_BF16X3_PACKED("sgemm_bf16x3d_mfma_256x256w8n", 256, 256, 2, 4, 463.0f, 1),
    ICL_SGEMM_BF16X3_PACKED("sgemm_bf16x3d_mfma_128x128", 128, 128, 2, 2, 366.0f, 2),
};
ICL_REGISTER_VARIANTS(kTable)
