// Grouped products of depth 256 (icl_panel_gemm) for the panel solves of the
// blocked TRSMs: the panel step [X_p; B_r] = [W_p; -M_p] B_p (+ [0; B_r]) and the
// batched M_p = L_rp W_p that precedes the steps.
//
// A GEMM of depth 256 is too short for the big GEMM kernels' tiles: at 128 x 128
// or 256 x 256 a 4096 x 1024 step has 256 or 64 tiles, a fraction of the GPU,
// and their slab pipelines barely fill before the K loop ends. Two kernels:
//
// panel_gemm_dma_kernel (used whenever the operands allow 16-byte copies): a
// workgroup owns a 64 x 64 tile of D and each of its four waves a 32 x 32 block
// over the full depth. The depth runs in eight 32-deep chunks copied global ->
// LDS with global_load_lds_dwordx4 (no VGPR staging), one chunk ahead of the
// MFMAs in a two-stage ring of 32 KB, so four or five workgroups fit on a CU.
// An operand whose rows (i for P, l for Q) are contiguous is stored S[k][64 rows]
// (one 1 KB copy per four k rows); one whose k is contiguous S[row][32 k], 128-byte
// rows of eight 16-byte chunks with logical chunk c of row r in slot c ^ ((r >> 1)
// & 7), so the 16-byte fragment reads of 16 consecutive rows are free of bank
// conflicts. Within a chunk, MFMA j of a wave takes k = 4 (2 (j / 4) + h) + j % 4
// for lane half h in both operands (a common permutation of the summation index),
// so one 16-byte read of a k-contiguous operand feeds four MFMAs. The MFMAs of a
// wave alternate between two accumulators, added before the store. D's old
// values (beta != 0) are loaded under the last chunk's MFMAs.
//
// panel_gemm_kernel (the fallback for unaligned operands): the four waves of a
// workgroup each take a 64-deep slice of the depth for the whole 64 x 64 tile,
// load it straight into the MFMA operand registers (one wait per tile) and add
// their partial tiles through LDS.
//
// The MFMA's lane dimension is the contiguous dimension of D, so D is written as
// 128-byte (DMA kernel) or 256-byte (fallback) column pieces.
#include "kernel_abi.h"
#include "mfma.hpp"

namespace {

constexpr int kPgT = 64;   // tile: 64 rows (i) x 64 lines (l)
constexpr int kPgKW = 64;  // depth per wave; four waves make 256
constexpr int kPgGroups = kPgKW / 8;

struct PgDims {
    int rows, lines, tile0;
};

// P_I: P contiguous along i (else along k); Q_K: Q contiguous along k (else along
// l); D_I: D contiguous along i (else along l). EDGE: some problem has rows or
// lines not a multiple of 64. VEC: every k-contiguous operand allows 16-byte
// loads (aligned base, stride a multiple of 4).
template <bool P_I, bool Q_K, bool D_I, bool EDGE, bool VEC>
__global__ void __launch_bounds__(256, 2) panel_gemm_kernel(icl_panel_gemm_args a)
{
    __shared__ float part[4 * kPgT * kPgT]; // the four waves' partial tiles, [w][u][v]
    const int b = blockIdx.x;
    int pi = 0;
    for (int i = 1; i < a.nprob; ++i)
        if (b >= a.prob[i].tile0) pi = i;
    const icl_panel_gemm_problem& pr = a.prob[pi];
    const int rows = pr.rows, lines = pr.lines;
    const int tiles_i = (rows + kPgT - 1) / kPgT;
    const int t = b - pr.tile0;
    const int i0 = (t % tiles_i) * kPgT, l0 = (t / tiles_i) * kPgT;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, c = lane & 31, h = lane >> 5;
    const int kw = w * kPgKW;

    // pv[q][s][j] = P(i0 + 32 j + c, k), qv[q][s][j] = Q(k, l0 + 32 j + c),
    // k = kw + 8 q + 4 h + s
    float pv[kPgGroups][4][2], qv[kPgGroups][4][2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int i = i0 + 32 * j + c, l = l0 + 32 * j + c;
        const bool iok = !EDGE || i < rows, lok = !EDGE || l < lines;
        const float* prow = pr.P + (long)(iok ? i : 0) * pr.p_is;
        const float* qcol = pr.Q + (long)(lok ? l : 0) * pr.q_ls;
#pragma unroll
        for (int q = 0; q < kPgGroups; ++q) {
            const int kb = kw + 8 * q + 4 * h;
            if constexpr (!P_I && VEC) {
                const f32x4 v = *reinterpret_cast<const f32x4*>(prow + kb);
#pragma unroll
                for (int s = 0; s < 4; ++s) pv[q][s][j] = iok ? v[s] : 0.0f;
            } else {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const float x = prow[(long)(kb + s) * pr.p_ks];
                    pv[q][s][j] = iok ? x : 0.0f;
                }
            }
            if constexpr (Q_K && VEC) {
                const f32x4 v = *reinterpret_cast<const f32x4*>(qcol + kb);
#pragma unroll
                for (int s = 0; s < 4; ++s) qv[q][s][j] = lok ? v[s] : 0.0f;
            } else {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const float x = qcol[(long)(kb + s) * pr.q_ks];
                    qv[q][s][j] = lok ? x : 0.0f;
                }
            }
        }
    }

    // Every load above is issued before the first MFMA: without this fence the
    // scheduler sinks each k-group's loads next to their MFMAs, and the wave then
    // waits out one memory latency per k-group instead of one per tile.
    __builtin_amdgcn_sched_barrier(0);

    // acc[r][v]: the MFMA's lanes run along D's contiguous dimension v (i when
    // D_I, l otherwise) and its registers along the other one, u.
    f32x16 acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int v = 0; v < 2; ++v) acc[r][v] = f32x16{};
#pragma unroll
    for (int q = 0; q < kPgGroups; ++q)
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    const float ua = D_I ? qv[q][s][r] : pv[q][s][r];
                    const float vb = D_I ? pv[q][s][v] : qv[q][s][v];
                    acc[r][v] = icl::mfma_32x32x2_f32(ua, vb, acc[r][v]);
                }

    float* mine = part + w * kPgT * kPgT;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int v = 0; v < 2; ++v)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int u = 32 * r + icl::mfma32_row(lane, e);
                mine[u * kPgT + 32 * v + c] = acc[r][v][e];
            }
    __syncthreads();

    // D = alpha * (sum of the four partials) + beta * D, one 256-byte row of the
    // contiguous dimension per wave and step.
    const int v = lane;
    const float alpha = pr.alpha, beta = pr.beta;
#pragma unroll 4
    for (int u = w; u < kPgT; u += 4) {
        const int o = u * kPgT + v;
        const float sum = part[o] + part[kPgT * kPgT + o] + part[2 * kPgT * kPgT + o] + part[3 * kPgT * kPgT + o];
        const int i = i0 + (D_I ? v : u), l = l0 + (D_I ? u : v);
        if (EDGE && (i >= rows || l >= lines)) continue;
        float* d = pr.D + (long)i * pr.d_is + (long)l * pr.d_ls;
        float x = alpha * sum;
        if (beta != 0.0f) x = fmaf(beta, *d, x);
        *d = x;
    }
}

// ---- LDS-DMA kernel ----
typedef __attribute__((address_space(1))) void pg_global_void;
typedef __attribute__((address_space(3))) void pg_lds_void;

constexpr int kPdKC = 32;                       // depth per chunk
constexpr int kPdChunks = 256 / kPdKC;
constexpr int kPdRowB = kPdKC * 4;              // bytes per row of a k-contiguous chunk
constexpr int kPdOpB = kPgT * kPdKC * 4;        // bytes of one operand chunk (8 KB)
constexpr int kPdStage = 2 * kPdOpB;            // P then Q
constexpr int kPdLds = 2 * kPdStage;            // two stages: 32 KB
constexpr int kPdBlocksPerWave = kPdOpB / 1024 / 4;

__device__ __forceinline__ int pd_swizzle(int r) { return (r >> 1) & 7; }

// Copies the 64-row x 32-deep chunk of an operand X(row, k) = X[row * rs + k * ks]
// (rs == 1 when ROWC, else ks == 1) at rows row0.., depth k0.. into S. Rows at or
// beyond `rows` read a valid row instead (their products are never stored); a
// ROWC copy moves four rows, so there rows % 4 == 0 (checked by the host).
template <bool ROWC>
__device__ __forceinline__ void pd_issue(const float* X, long rs, long ks, int row0, int k0, int rows,
                                         unsigned char* S, int wave, int lane)
{
#pragma unroll
    for (int i = 0; i < kPdBlocksPerWave; ++i) {
        const int blk = wave + 4 * i; // 1 KB block of the chunk
        const float* g;
        if constexpr (ROWC) {
            const int kr = 4 * blk + (lane >> 4); // k row; lane covers rows 4 (lane % 16) .. + 3
            const int r = min(row0 + 4 * (lane & 15), rows - 4);
            g = X + (long)(k0 + kr) * ks + r;
        } else {
            const int r = 8 * blk + (lane >> 3);            // row of the chunk
            const int c = (lane & 7) ^ pd_swizzle(r);       // logical 16-byte chunk in this slot
            g = X + (long)min(row0 + r, rows - 1) * rs + k0 + 4 * c;
        }
        __builtin_amdgcn_global_load_lds((pg_global_void*)g, (pg_lds_void*)(S + blk * 1024), 16, 0, 0);
    }
}

// f[t] = X(rb + lane % 32, k) for MFMA 4 g + t of the chunk, k = 4 (2 g + h) + t.
template <bool ROWC>
__device__ __forceinline__ void pd_frag(const unsigned char* S, int rb, int g, int lane, float (&f)[4])
{
    const int h = lane >> 5, r = rb + (lane & 31), q = 2 * g + h;
    if constexpr (ROWC) {
        const float* s = reinterpret_cast<const float*>(S);
#pragma unroll
        for (int t = 0; t < 4; ++t) f[t] = s[(4 * q + t) * kPgT + r];
    } else {
        const f32x4 v = *reinterpret_cast<const f32x4*>(S + r * kPdRowB + ((q ^ pd_swizzle(r)) << 4));
#pragma unroll
        for (int t = 0; t < 4; ++t) f[t] = v[t];
    }
}

template <bool P_I, bool Q_K, bool D_I, bool EDGE>
__global__ void __launch_bounds__(256, 2) panel_gemm_dma_kernel(icl_panel_gemm_args a)
{
    extern __shared__ __attribute__((aligned(1024))) unsigned char pd_lds[];
    const int b = blockIdx.x;
    int pi = 0;
    for (int i = 1; i < a.nprob; ++i)
        if (b >= a.prob[i].tile0) pi = i;
    const icl_panel_gemm_problem& pr = a.prob[pi];
    const int rows = pr.rows, lines = pr.lines;
    const int tiles_i = (rows + kPgT - 1) / kPgT;
    const int t = b - pr.tile0;
    const int i0 = (t % tiles_i) * kPgT, l0 = (t / tiles_i) * kPgT;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wi = 32 * (wave & 1), wl = 32 * (wave >> 1);

    auto issue = [&](int c, int s) {
        unsigned char* st = pd_lds + s * kPdStage;
        pd_issue<P_I>(pr.P, pr.p_is, pr.p_ks, i0, c * kPdKC, EDGE ? rows : 1 << 30, st, wave, lane);
        pd_issue<!Q_K>(pr.Q, pr.q_ls, pr.q_ks, l0, c * kPdKC, EDGE ? lines : 1 << 30, st + kPdOpB, wave, lane);
    };

    // Element e of a lane's 32 x 32 block: D(i, l) at (lane % 32, row e) or the other way round.
    auto d_index = [&](int e, int& i, int& l) {
        const int u = icl::mfma32_row(lane, e), v = lane & 31;
        i = i0 + wi + (D_I ? v : u);
        l = l0 + wl + (D_I ? u : v);
    };
    const float alpha = pr.alpha, beta = pr.beta;
    float dold[16];

    f32x16 acc[2] = {f32x16{}, f32x16{}};
    issue(0, 0);
#pragma unroll
    for (int c = 0; c < kPdChunks; ++c) {
        // chunk c in LDS for every wave, and every wave done with chunk c - 1
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (c + 1 < kPdChunks) issue(c + 1, (c + 1) & 1);
        if (c + 1 == kPdChunks && beta != 0.0f) {
            // D's old values are loaded under the last chunk's MFMAs rather than after them
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                int i, l;
                d_index(e, i, l);
                dold[e] = (!EDGE || (i < rows && l < lines)) ? pr.D[(long)i * pr.d_is + (long)l * pr.d_ls] : 0.0f;
            }
        }
        const unsigned char* Ps = pd_lds + (c & 1) * kPdStage;
        const unsigned char* Qs = Ps + kPdOpB;
#pragma unroll
        for (int g = 0; g < kPdKC / 8; ++g) {
            float pf[4], qf[4];
            pd_frag<P_I>(Ps, wi, g, lane, pf);
            pd_frag<!Q_K>(Qs, wl, g, lane, qf);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                f32x16& ac = acc[s & 1];
                // lanes along D's contiguous dimension: the b operand supplies it
                ac = D_I ? icl::mfma_32x32x2_f32(qf[s], pf[s], ac) : icl::mfma_32x32x2_f32(pf[s], qf[s], ac);
            }
        }
    }
    const f32x16 sum = acc[0] + acc[1];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        int i, l;
        d_index(e, i, l);
        if (EDGE && (i >= rows || l >= lines)) continue;
        float x = alpha * sum[e];
        if (beta != 0.0f) x = fmaf(beta, dold[e], x);
        pr.D[(long)i * pr.d_is + (long)l * pr.d_ls] = x;
    }
}

// ---- complex LDS-DMA kernel (interleaved re, im) ----
// The same tiling as panel_gemm_dma_kernel with complex elements: 16-deep
// chunks of 64 rows (8 KB per operand, as for the real kernel). Rows-contiguous
// operands are stored S[k][64 rows] (512-byte k rows, one 1 KB copy per two k
// rows); k-contiguous ones S[row][16 k], 128-byte rows of eight 16-byte chunks
// of two complex k, chunk c of row r in slot c ^ ((r >> 1) & 7). Step j = 2 g + s
// of a chunk takes complex k = 2 (2 g + h) + s for lane half h in both operands,
// so one 16-byte read of a k-contiguous operand feeds two steps; a step is four
// real MFMAs: re += Pr Qr - Pi Qi, im += Pr Qi + Pi Qr (conjugation negates Pi or Qi).
constexpr int kCdKC = 16;
constexpr int kCdChunks = 256 / kCdKC;

// rows % 2 == 0 for ROWC (each copy moves two complex rows), checked by the host.
template <bool ROWC>
__device__ __forceinline__ void cd_issue(const float* X, long rs, long ks, int row0, int k0, int rows,
                                         unsigned char* S, int wave, int lane)
{
#pragma unroll
    for (int i = 0; i < kPdBlocksPerWave; ++i) {
        const int blk = wave + 4 * i;
        const float* g;
        if constexpr (ROWC) {
            const int kr = 2 * blk + (lane >> 5);                  // k row; lane covers rows 2 (lane % 32), +1
            const int r = min(row0 + 2 * (lane & 31), rows - 2);
            g = X + 2 * ((long)(k0 + kr) * ks + r);
        } else {
            const int r = 8 * blk + (lane >> 3);
            const int c = (lane & 7) ^ pd_swizzle(r);
            g = X + 2 * ((long)min(row0 + r, rows - 1) * rs + k0 + 2 * c);
        }
        __builtin_amdgcn_global_load_lds((pg_global_void*)g, (pg_lds_void*)(S + blk * 1024), 16, 0, 0);
    }
}

// re[s], im[s] = X(rb + lane % 32, k), k = 2 (2 g + h) + s, s = 0, 1.
template <bool ROWC>
__device__ __forceinline__ void cd_frag(const unsigned char* S, int rb, int g, int lane, float (&re)[2],
                                        float (&im)[2])
{
    const int h = lane >> 5, r = rb + (lane & 31), q = 2 * g + h;
    if constexpr (ROWC) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const f32x2 v = *reinterpret_cast<const f32x2*>(S + (2 * q + s) * 512 + r * 8);
            re[s] = v[0];
            im[s] = v[1];
        }
    } else {
        const f32x4 v = *reinterpret_cast<const f32x4*>(S + r * 128 + ((q ^ pd_swizzle(r)) << 4));
        re[0] = v[0], im[0] = v[1], re[1] = v[2], im[1] = v[3];
    }
}

template <bool P_I, bool Q_K, bool D_I, bool EDGE>
__global__ void __launch_bounds__(256, 2) cpanel_gemm_dma_kernel(icl_panel_gemm_args a)
{
    extern __shared__ __attribute__((aligned(1024))) unsigned char cd_lds[];
    const int b = blockIdx.x;
    int pi = 0;
    for (int i = 1; i < a.nprob; ++i)
        if (b >= a.prob[i].tile0) pi = i;
    const icl_panel_gemm_problem& pr = a.prob[pi];
    const int rows = pr.rows, lines = pr.lines;
    const int tiles_i = (rows + kPgT - 1) / kPgT;
    const int t = b - pr.tile0;
    const int i0 = (t % tiles_i) * kPgT, l0 = (t / tiles_i) * kPgT;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wi = 32 * (wave & 1), wl = 32 * (wave >> 1);
    const float sp = pr.conj_p ? -1.0f : 1.0f, sq = pr.conj_q ? -1.0f : 1.0f;

    auto issue = [&](int c, int s) {
        unsigned char* st = cd_lds + s * kPdStage;
        cd_issue<P_I>(pr.P, pr.p_is, pr.p_ks, i0, c * kCdKC, EDGE ? rows : 1 << 30, st, wave, lane);
        cd_issue<!Q_K>(pr.Q, pr.q_ls, pr.q_ks, l0, c * kCdKC, EDGE ? lines : 1 << 30, st + kPdOpB, wave, lane);
    };
    auto d_index = [&](int e, int& i, int& l) {
        const int u = icl::mfma32_row(lane, e), v = lane & 31;
        i = i0 + wi + (D_I ? v : u);
        l = l0 + wl + (D_I ? u : v);
    };
    const float alpha = pr.alpha, beta = pr.beta;
    f32x2 dold[16];

    f32x16 ar = f32x16{}, ai = f32x16{};
    issue(0, 0);
    for (int c = 0; c < kCdChunks; ++c) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (c + 1 < kCdChunks) issue(c + 1, (c + 1) & 1);
        if (c + 1 == kCdChunks && beta != 0.0f) {
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                int i, l;
                d_index(e, i, l);
                dold[e] = (!EDGE || (i < rows && l < lines))
                              ? *reinterpret_cast<const f32x2*>(pr.D + 2 * ((long)i * pr.d_is + (long)l * pr.d_ls))
                              : f32x2{0.0f, 0.0f};
            }
        }
        const unsigned char* Ps = cd_lds + (c & 1) * kPdStage;
        const unsigned char* Qs = Ps + kPdOpB;
#pragma unroll
        for (int g = 0; g < kCdKC / 4; ++g) {
            float pr_[2], pi_[2], qr[2], qi[2];
            cd_frag<P_I>(Ps, wi, g, lane, pr_, pi_);
            cd_frag<!Q_K>(Qs, wl, g, lane, qr, qi);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const float p_im = sp * pi_[s], q_im = sq * qi[s];
                if constexpr (D_I) { // b operand: P (lanes along i)
                    ar = icl::mfma_32x32x2_f32(qr[s], pr_[s], ar);
                    ar = icl::mfma_32x32x2_f32(-q_im, p_im, ar);
                    ai = icl::mfma_32x32x2_f32(qr[s], p_im, ai);
                    ai = icl::mfma_32x32x2_f32(q_im, pr_[s], ai);
                } else {
                    ar = icl::mfma_32x32x2_f32(pr_[s], qr[s], ar);
                    ar = icl::mfma_32x32x2_f32(-p_im, q_im, ar);
                    ai = icl::mfma_32x32x2_f32(pr_[s], q_im, ai);
                    ai = icl::mfma_32x32x2_f32(p_im, qr[s], ai);
                }
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        int i, l;
        d_index(e, i, l);
        if (EDGE && (i >= rows || l >= lines)) continue;
        float xr = alpha * ar[e], xi = alpha * ai[e];
        if (beta != 0.0f) {
            xr = fmaf(beta, dold[e][0], xr);
            xi = fmaf(beta, dold[e][1], xi);
        }
        *reinterpret_cast<f32x2*>(pr.D + 2 * ((long)i * pr.d_is + (long)l * pr.d_ls)) = f32x2{xr, xi};
    }
}

using PgKern = void (*)(icl_panel_gemm_args);

template <bool P_I, bool Q_K, bool D_I>
PgKern pick(bool edge, bool vec, bool dma)
{
    if (dma) return edge ? panel_gemm_dma_kernel<P_I, Q_K, D_I, true> : panel_gemm_dma_kernel<P_I, Q_K, D_I, false>;
    if (edge) return vec ? panel_gemm_kernel<P_I, Q_K, D_I, true, true> : panel_gemm_kernel<P_I, Q_K, D_I, true, false>;
    return vec ? panel_gemm_kernel<P_I, Q_K, D_I, false, true> : panel_gemm_kernel<P_I, Q_K, D_I, false, false>;
}

bool aligned16(const void* p, long stride) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0 && stride % 4 == 0; }

} // namespace

extern "C" hipError_t icl_panel_gemm(icl_panel_gemm_args* a, hipStream_t stream)
{
    if (a->nprob <= 0) return hipSuccess;
    if (a->nprob > ICL_PANEL_GEMM_MAX) return hipErrorInvalidValue;
    const icl_panel_gemm_problem& f = a->prob[0];
    const bool p_i = f.p_is == 1, q_k = f.q_ks == 1, d_i = f.d_is == 1;
    long tiles = 0;
    bool edge = false, vec = true;
    // The DMA kernel copies 16-byte pieces: every operand 16-byte aligned with its
    // non-unit stride a multiple of 4 floats, and the row count of an operand whose
    // rows are contiguous a multiple of 4.
    bool dma = true;
    for (int i = 0; i < a->nprob; ++i) {
        icl_panel_gemm_problem& p = a->prob[i];
        if (p.rows < 0 || p.lines < 0) return hipErrorInvalidValue;
        // the layout is one per launch
        if ((p.p_is == 1) != p_i || (!p_i && p.p_ks != 1) || (p.q_ks == 1) != q_k || (!q_k && p.q_ls != 1) ||
            (p.d_is == 1) != d_i || (!d_i && p.d_ls != 1))
            return hipErrorInvalidValue;
        p.tile0 = (int)tiles;
        tiles += (long)((p.rows + kPgT - 1) / kPgT) * ((p.lines + kPgT - 1) / kPgT);
        edge = edge || p.rows % kPgT != 0 || p.lines % kPgT != 0;
        if (!p_i) vec = vec && aligned16(p.P, p.p_is);
        if (q_k) vec = vec && aligned16(p.Q, p.q_ls);
        dma = dma && aligned16(p.P, p_i ? p.p_ks : p.p_is) && aligned16(p.Q, q_k ? p.q_ls : p.q_ks) &&
              (!p_i || p.rows % 4 == 0) && (q_k || p.lines % 4 == 0);
    }
    if (tiles == 0) return hipSuccess;
    if (tiles > 0x7fffffffL) return hipErrorInvalidValue;
    PgKern k = nullptr;
    if (p_i && q_k && d_i) k = pick<true, true, true>(edge, vec, dma);          // left steps; M_p of op N / right op N
    else if (!p_i && !q_k && !d_i) k = pick<false, false, false>(edge, vec, dma); // right steps
    else if (!p_i && q_k && d_i) k = pick<false, true, true>(edge, vec, dma);   // M_p, left op T
    else if (p_i && !q_k && d_i) k = pick<true, false, true>(edge, vec, dma);   // N_p, right op T
    else return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3((unsigned)tiles), dim3(256), dma ? kPdLds : 0, stream, *a);
    return hipGetLastError();
}

namespace {

// Layout of a complex launch: the contiguous dimensions, alike in all problems,
// and whether every operand allows the 16-byte copies.
bool cpanel_layout(const icl_panel_gemm_args* a, bool* p_i, bool* q_k, bool* d_i)
{
    if (a->nprob <= 0 || a->nprob > ICL_PANEL_GEMM_MAX) return false;
    const icl_panel_gemm_problem& f = a->prob[0];
    *p_i = f.p_is == 1, *q_k = f.q_ks == 1, *d_i = f.d_is == 1;
    for (int i = 0; i < a->nprob; ++i) {
        const icl_panel_gemm_problem& p = a->prob[i];
        if (p.rows < 0 || p.lines < 0) return false;
        if ((p.p_is == 1) != *p_i || (!*p_i && p.p_ks != 1) || (p.q_ks == 1) != *q_k || (!*q_k && p.q_ls != 1) ||
            (p.d_is == 1) != *d_i || (!*d_i && p.d_ls != 1))
            return false;
        // 16-byte copies: aligned bases, even non-unit strides (two complex per 16 bytes)
        auto ok = [](const void* x, long stride) { return (reinterpret_cast<uintptr_t>(x) & 15u) == 0 && stride % 2 == 0; };
        if (!ok(p.P, *p_i ? p.p_ks : p.p_is) || !ok(p.Q, *q_k ? p.q_ls : p.q_ks)) return false;
        if ((*p_i && p.rows % 2 != 0) || (!*q_k && p.lines % 2 != 0)) return false;
    }
    const bool supported = (*p_i && *q_k && *d_i) || (!*p_i && *q_k && *d_i) || (*p_i && !*q_k && *d_i) ||
                           (!*p_i && !*q_k && !*d_i);
    return supported;
}

template <bool P_I, bool Q_K, bool D_I>
PgKern cpick(bool edge)
{
    return edge ? cpanel_gemm_dma_kernel<P_I, Q_K, D_I, true> : cpanel_gemm_dma_kernel<P_I, Q_K, D_I, false>;
}

} // namespace

extern "C" int icl_cpanel_gemm_ok(const icl_panel_gemm_args* a)
{
    bool p_i, q_k, d_i;
    return cpanel_layout(a, &p_i, &q_k, &d_i) ? 1 : 0;
}

extern "C" hipError_t icl_cpanel_gemm(icl_panel_gemm_args* a, hipStream_t stream)
{
    if (a->nprob <= 0) return hipSuccess;
    bool p_i, q_k, d_i;
    if (!cpanel_layout(a, &p_i, &q_k, &d_i)) return hipErrorInvalidValue;
    long tiles = 0;
    bool edge = false;
    for (int i = 0; i < a->nprob; ++i) {
        icl_panel_gemm_problem& p = a->prob[i];
        p.tile0 = (int)tiles;
        tiles += (long)((p.rows + kPgT - 1) / kPgT) * ((p.lines + kPgT - 1) / kPgT);
        edge = edge || p.rows % kPgT != 0 || p.lines % kPgT != 0;
    }
    if (tiles == 0) return hipSuccess;
    if (tiles > 0x7fffffffL) return hipErrorInvalidValue;
    PgKern k = nullptr;
    if (p_i && q_k && d_i) k = cpick<true, true, true>(edge);           // left op N, right op N
    else if (!p_i && q_k && d_i) k = cpick<false, true, true>(edge);    // left op T / C
    else if (p_i && !q_k && d_i) k = cpick<true, false, true>(edge);    // right op T / C
    else k = cpick<false, false, false>(edge);
    hipLaunchKernelGGL(k, dim3((unsigned)tiles), dim3(256), kPdLds, stream, *a);
    return hipGetLastError();
}
