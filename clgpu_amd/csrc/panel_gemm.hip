// Grouped products of depth 256 (icl_panel_gemm) for the panel solves of the
// blocked TRSMs: the panel step [X_p; B_r] = [W_p; -M_p] B_p (+ [0; B_r]) and the
// batched M_p = L_rp W_p that precedes the steps.
//
// A GEMM of depth 256 is too short for a slab pipeline: with one 16-deep slab
// in flight, a 64 x 64 tile waits out a global-memory latency per slab and the
// launch takes 10-30 us whatever its size. Here a workgroup owns a 64 x 64 tile
// of D and its four waves each take a 64-deep slice of the depth; a wave issues
// every load of its slice at once, straight into the MFMA operand registers
// (one wait per tile), runs 128 v_mfma_f32_32x32x2_f32, and the four partial
// tiles are added through LDS in the epilogue. Two workgroups fit on a CU.
//
// Operand fragments: lane l (c = l % 32, h = l / 32) supplies, for MFMA s of
// k-group q, the value at k = 64 w + 8 q + 4 h + s -- the same permutation of
// the summation index for both operands -- so an operand whose k is contiguous
// is read as one 16-byte load per (q, 32-block), and an operand whose other
// dimension is contiguous as 4-byte loads coalesced over c.
// The MFMA's lane dimension is the contiguous dimension of D, so the partial
// tiles are written to LDS and D is written as 256-byte rows.
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

using PgKern = void (*)(icl_panel_gemm_args);

template <bool P_I, bool Q_K, bool D_I>
PgKern pick(bool edge, bool vec)
{
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
    }
    if (tiles == 0) return hipSuccess;
    if (tiles > 0x7fffffffL) return hipErrorInvalidValue;
    PgKern k = nullptr;
    if (p_i && q_k && d_i) k = pick<true, true, true>(edge, vec);          // left steps; M_p of op N / right op N
    else if (!p_i && !q_k && !d_i) k = pick<false, false, false>(edge, vec); // right steps
    else if (!p_i && q_k && d_i) k = pick<false, true, true>(edge, vec);   // M_p, left op T
    else if (p_i && !q_k && d_i) k = pick<true, false, true>(edge, vec);   // N_p, right op T
    else return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3((unsigned)tiles), dim3(256), 0, stream, *a);
    return hipGetLastError();
}
