// STRSM building blocks for gfx950: the diagonal-block solver and inverter,
// the blocked / dataflow / panel solves, the FP64 residual and stopping test
// of iterative refinement, and 2-D elementwise updates.
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
    // Sixteen loads in flight at a time (an entry outside the strict lower
    // triangle reads the block's first diagonal element instead and stores 0).
#pragma unroll
    for (int i0 = 0; i0 < kNB; i0 += 16) {
        float v[16];
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
            const int i = i0 + ii;
            const bool used = i < nb && lane < nb && lane > i;
            v[ii] = L[used ? (long)lane * l_si + (long)i * l_sk : 0];
        }
#pragma unroll
        for (int ii = 0; ii < 16; ++ii) {
            const int i = i0 + ii;
            const bool used = i < nb && lane < nb && lane > i;
            s.Lt[i][lane] = used ? v[ii] : 0.0f;
        }
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
// (workgroup b handles rows/columns [64 b, 64 b + nb)) and writes inv(L_bb) to
// invD[b * 4096 + j * 64 + i] (column-major, rows and columns beyond nb zero),
// by recursive doubling: each wave inverts one 16 x 16 diagonal sub-block by
// substitution (lane j forms column j, the sub-block's entries read from LDS,
// no barrier per pivot), then the 32 x 32 halves and the whole block follow from
// inv([A 0; C D]) = [inv(A) 0; -inv(D) C inv(A), inv(D)], each a pair of small
// products spread over the workgroup. Rows beyond nb are an identity block, so
// the padded inverse is exact there and those entries are written as zeros.
constexpr int kIvLd = kNB + 1;
__global__ void __launch_bounds__(256) trsm_invert_kernel(icl_trsm_blocked_args a)
{
    __shared__ float Ls[kNB][kIvLd];  // L(i, k), i >= k; identity beyond nb
    __shared__ float Is[kNB][kIvLd];  // the inverse, built in place
    __shared__ float Ts[kNB / 2][kIvLd]; // products' intermediate
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b = blockIdx.x;
    const int r0 = b * kNB;
    const int nb = min(kNB, a.nsys - r0);
    const float* L = a.L + (long)r0 * (a.l_si + a.l_sk);
    {
        // 16 entries per thread, all loads in flight together: entry e = tid + 256 q
        // is (i, k) = (e / 64, e % 64), so a wave reads one row.
        // Each entry, the diagonal included, has exactly one writer here.
        float v[16];
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
            Ls[i][k] = used ? v[q] : (k == i ? 1.0f : 0.0f);
            Is[i][k] = 0.0f;
        }
    }
    __syncthreads();
    // 16 x 16 diagonal sub-block w: lane j < 16 forms column j of its inverse.
    {
        const int o = 16 * wave, j = lane;
        if (j < 16) {
            float x[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                float acc = i == j ? 1.0f : 0.0f;
#pragma unroll
                for (int k = 0; k < i; ++k) acc -= Ls[o + i][o + k] * x[k];
                x[i] = acc / Ls[o + i][o + i];
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
            float out[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q >= per) break;
                const int e = tid + 256 * q;
                const int blk = e / (h * h), rc = e % (h * h), r = rc / h, c = rc % h;
                const int o = blk * 2 * h;
                float acc = 0.0f;
                if (step == 0) { // T(r, c) = sum_k C(r, k) inv(A)(k, c), k >= c
                    for (int k = c; k < h; ++k) acc += Ls[o + h + r][o + k] * Is[o + k][o + c];
                } else { // -sum_k inv(D)(r, k) T(k, c), k <= r
                    for (int k = 0; k <= r; ++k) acc -= Is[o + h + r][o + h + k] * Ts[blk * h + k][c];
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
    float* dst = a.invD + (long)b * kNB * kNB;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int e = tid + 256 * q, j = e / kNB, i = e % kNB; // column j, row i: consecutive threads along i
        dst[e] = (i < nb && j < nb && i >= j) ? Is[i][j] : 0.0f;
    }
    if (a.sync && b == 0) { // the dataflow solve that follows on this stream starts from zeroed flags
        const int n = ((a.nsys + kNB - 1) / kNB) * ((a.nlines + kNB - 1) / kNB) + 2;
        for (int i = tid; i < n; i += 256) a.sync[i] = 0;
    }
}

// ---------------------------------------------------------------------------
// FP64 residual R = fp32(alpha * B0 - P * Q) on v_mfma_f64_16x16x4_f64.
// 64 x 64 output tiles, 4 waves per workgroup, each 32 x 32 = 2 x 2 MFMA blocks.
// K slabs of 32 are staged in LDS as FP32 (converted to FP64 at the fragment
// read) in two buffers: the next slab is read into registers while the current
// one is multiplied and stored to the other buffer afterwards, one barrier per
// slab. Slabs outside the triangle are skipped and only slabs that cross the
// diagonal apply the triangle rule. The K range of a tile depends on its
// position along the triangular operand's free dimension, so a workgroup takes
// tiles t and T - 1 - t there (T tiles): every workgroup then does the same
// total K, and no CU is left with the long tiles while others idle.
// ---------------------------------------------------------------------------
constexpr int RT = 64, RK = 32;
constexpr int kResLd = RT + 16;        // the four 16-lane row groups of a fragment read hit disjoint banks

// One thread's share of a slab of an operand whose element (i, k) (i < RT along
// the tile, k < RK along K) sits at X[i * s_i + k * s_k]: eight elements as two
// runs h of four e along the unit-stride dimension, so that an interior slab of
// a 16-byte aligned operand is read with two 16-byte loads. KFAST (s_k == 1):
// (i0 + 32 h, k0 + e); otherwise (s_i == 1): (i0 + e, k0 + 16 h).
template <bool KFAST>
struct ResMap {
    int i0, k0;
    __device__ __forceinline__ explicit ResMap(int tid)
    {
        i0 = KFAST ? tid / 8 : (tid % 16) * 4;
        k0 = KFAST ? (tid % 8) * 4 : tid / 16;
    }
    __device__ __forceinline__ int i(int h, int e) const { return KFAST ? i0 + 32 * h : i0 + e; }
    __device__ __forceinline__ int k(int h, int e) const { return KFAST ? k0 + e : k0 + 16 * h; }
};

template <bool PT, bool QT>
__global__ void __launch_bounds__(256, 3) residual_f64_kernel(icl_residual_args a)
{
    __shared__ __attribute__((aligned(16))) float Ps[2][RK][kResLd]; // Ps[b][k][r] = P(m0 + r, kb + k)
    __shared__ __attribute__((aligned(16))) float Qs[2][RK][kResLd]; // Qs[b][k][c] = Q(kb + k, n0 + c)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tiles_m = (a.M + RT - 1) / RT, tiles_n = (a.N + RT - 1) / RT;
    const int bid = blockIdx.x;
    // This workgroup's tiles: (tm[t], tn[t]) for t < ntiles.
    int tm[2], tn[2], ntiles = 1;
    if (a.tri_on == 1) { // the K range depends on the tile row
        tn[0] = tn[1] = bid % tiles_n;
        tm[0] = bid / tiles_n;
        tm[1] = tiles_m - 1 - tm[0];
        ntiles = tm[1] > tm[0] ? 2 : 1;
    } else if (a.tri_on == 2) { // on the tile column
        tm[0] = tm[1] = bid % tiles_m;
        tn[0] = bid / tiles_m;
        tn[1] = tiles_n - 1 - tn[0];
        ntiles = tn[1] > tn[0] ? 2 : 1;
    } else {
        tm[0] = tm[1] = bid % tiles_m;
        tn[0] = tn[1] = bid / tiles_m;
    }
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    // P(i, k) = P[i * ldp + k] when PT, P[i + k * ldp] otherwise; likewise Q(k, c)
    // = Q[k * ldq + c] when QT, Q[c * ldq + k] otherwise (c plays the role of i).
    const ResMap<PT> pm(tid);
    const ResMap<!QT> qm(tid);
    const long p_si = PT ? a.ldp : 1, p_sk = PT ? 1 : a.ldp;
    const long q_si = QT ? 1 : a.ldq, q_sk = QT ? a.ldq : 1;
    const bool p_vec = (reinterpret_cast<uintptr_t>(a.P) & 15) == 0 && (a.ldp & 3) == 0;
    const bool q_vec = (reinterpret_cast<uintptr_t>(a.Q) & 15) == 0 && (a.ldq & 3) == 0;

    for (int t = 0; t < ntiles; ++t) {
        const int m0 = tm[t] * RT, n0 = tn[t] * RT;
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

        // fetch() only loads (an element outside the matrices reads element 0
        // instead), so that all loads of a slab are in flight together; stash()
        // applies the bounds and the triangle rule when it stores to LDS.
        float pv[2][4], qv[2][4];
        auto fetch = [&](int kb) {
            const float* pb = a.P + (long)m0 * p_si + (long)kb * p_sk; // P(m0, kb)
            const float* qb = a.Q + (long)n0 * q_si + (long)kb * q_sk; // Q(kb, n0)
            const bool k_in = kb + RK <= a.K;
            if (p_vec && k_in && m0 + RT <= a.M) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const float4 t = *reinterpret_cast<const float4*>(pb + (long)pm.i(h, 0) * p_si +
                                                                      (long)pm.k(h, 0) * p_sk);
                    pv[h][0] = t.x, pv[h][1] = t.y, pv[h][2] = t.z, pv[h][3] = t.w;
                }
            } else {
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int i = pm.i(h, e), k = pm.k(h, e);
                        const bool ok = m0 + i < a.M && kb + k < a.K;
                        pv[h][e] = *(ok ? pb + (long)i * p_si + (long)k * p_sk : a.P);
                    }
            }
            if (q_vec && k_in && n0 + RT <= a.N) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const float4 t = *reinterpret_cast<const float4*>(qb + (long)qm.i(h, 0) * q_si +
                                                                      (long)qm.k(h, 0) * q_sk);
                    qv[h][0] = t.x, qv[h][1] = t.y, qv[h][2] = t.z, qv[h][3] = t.w;
                }
            } else {
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int c = qm.i(h, e), k = qm.k(h, e);
                        const bool ok = n0 + c < a.N && kb + k < a.K;
                        qv[h][e] = *(ok ? qb + (long)c * q_si + (long)k * q_sk : a.Q);
                    }
            }
        };
        auto stash = [&](int b, int kb) {
            // triangle rule only where the slab crosses the diagonal of the triangular operand
            const bool p_diag = a.tri_on == 1 && kb < m0 + RT && m0 < kb + RK;
            const bool q_diag = a.tri_on == 2 && kb < n0 + RT && n0 < kb + RK;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int i = m0 + pm.i(h, e), k = kb + pm.k(h, e);
                    float x = (i < a.M && k < a.K) ? pv[h][e] : 0.0f;
                    x = (p_diag && (a.tri_lower ? k > i : k < i)) ? 0.0f : x;
                    v[e] = (p_diag && a.tri_unit && k == i) ? 1.0f : x;
                }
                if (PT) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) Ps[b][pm.k(h, e)][pm.i(h, 0)] = v[e];
                } else {
                    *reinterpret_cast<float4*>(&Ps[b][pm.k(h, 0)][pm.i(h, 0)]) = make_float4(v[0], v[1], v[2], v[3]);
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) { // Q(k, c): row k, column c of the triangle
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int c = n0 + qm.i(h, e), k = kb + qm.k(h, e);
                    float x = (c < a.N && k < a.K) ? qv[h][e] : 0.0f;
                    x = (q_diag && (a.tri_lower ? c > k : c < k)) ? 0.0f : x;
                    v[e] = (q_diag && a.tri_unit && k == c) ? 1.0f : x;
                }
                if (QT) {
                    *reinterpret_cast<float4*>(&Qs[b][qm.k(h, 0)][qm.i(h, 0)]) = make_float4(v[0], v[1], v[2], v[3]);
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) Qs[b][qm.k(h, e)][qm.i(h, 0)] = v[e];
                }
            }
        };
        if (k_lo < k_hi) {
            fetch(k_lo);
            stash(0, k_lo);
            __syncthreads();
        }
        int cur = 0;
        for (int kb = k_lo; kb < k_hi; kb += RK) {
            const bool more = kb + RK < k_hi;
            if (more) fetch(kb + RK); // in flight during the products below
#pragma unroll
            for (int kk = 0; kk < RK; kk += 4) {
                const int kr = kk + (lane >> 4);
                double p[2], q[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) p[i] = (double)Ps[cur][kr][wm + 16 * i + (lane & 15)];
#pragma unroll
                for (int j = 0; j < 2; ++j) q[j] = (double)Qs[cur][kr][wn + 16 * j + (lane & 15)];
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) acc[i][j] = icl::mfma_16x16x4_f64(q[j], p[i], acc[i][j]);
            }
            // The other buffer was last read before the previous barrier.
            if (more) stash(cur ^ 1, kb + RK);
            __syncthreads();
            cur ^= 1;
        }

        // B0 for the whole fragment first (outside the matrix: element 0), then the stores.
        const int m_base = m0 + wm + (lane & 15), n_base = n0 + wn + (lane >> 4);
        float b0[2][2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m_base + 16 * i, n = n_base + 16 * j + 4 * r;
                    b0[i][j][r] = a.B0[(m < a.M && n < a.N) ? (long)n * a.ldb0 + m : 0];
                }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m_base + 16 * i, n = n_base + 16 * j + 4 * r;
                    if (m < a.M && n < a.N)
                        a.R[(long)n * a.ldr + m] = (float)(a.alpha * (double)b0[i][j][r] - acc[i][j][r]);
                }
    }
}

// ---------------------------------------------------------------------------
// The same residual with its slabs copied global -> LDS by DMA
// (global_load_lds_dwordx4): no VGPR staging, the next slab in flight while the
// current one is multiplied, the triangle rule applied at the fragment read of
// the slabs that cross the diagonal, and B0 loaded under the last slab's MFMAs.
// For problems without edges (M, N multiples of 64, K of 32) whose operands
// allow 16-byte copies; residual_f64_kernel runs the others.
// LDS layouts of a 64-row x 32-deep FP32 slab (rows: m for P, n for Q):
//  * rows contiguous: S[k][64], the rows of k row kr rotated by 16 ((kr >> 2) & 3)
//    positions, so the four k rows one fragment read touches (k = kk + 4 g + s,
//    g = lane / 16) sit on disjoint banks;
//  * k contiguous: S[row][32], 128-byte rows of eight 16-byte chunks, logical
//    chunk c of row r in slot c ^ ((r >> 1) & 7); a 16-byte read gives the four
//    k of one lane's MFMAs s = 0..3.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(1))) void rd_global_void;
typedef __attribute__((address_space(3))) void rd_lds_void;
constexpr int kRdOpB = RT * RK * 4; // 8 KB per operand slab
constexpr int kRdStage = 2 * kRdOpB;
constexpr int kRdLds = 2 * kRdStage; // two stages: 32 KB

__device__ __forceinline__ int rd_rot(int kr) { return 16 * ((kr >> 2) & 3); }

// X(row, k) = X[row * rs + k * ks], rs == 1 when ROWC, else ks == 1.
template <bool ROWC>
__device__ __forceinline__ void rd_issue(const float* X, long rs, long ks, int row0, int k0, unsigned char* S,
                                         int wave, int lane)
{
#pragma unroll
    for (int i = 0; i < kRdOpB / 1024 / 4; ++i) {
        const int blk = wave + 4 * i;
        const float* g;
        if constexpr (ROWC) {
            const int kr = 4 * blk + (lane >> 4);                  // k row
            const int r = (4 * (lane & 15) - rd_rot(kr)) & (RT - 1); // rows r..r+3 go to slot lane % 16
            g = X + (long)(k0 + kr) * ks + row0 + r;
        } else {
            const int r = 8 * blk + (lane >> 3);
            const int c = (lane & 7) ^ ((r >> 1) & 7);
            g = X + (long)(row0 + r) * rs + k0 + 4 * c;
        }
        __builtin_amdgcn_global_load_lds((rd_global_void*)g, (rd_lds_void*)(S + blk * 1024), 16, 0, 0);
    }
}

// v[s] = X(x, kk + 4 g + s) for s = 0..3, g = lane / 16.
template <bool ROWC>
__device__ __forceinline__ void rd_frag(const unsigned char* S, int x, int kk, int lane, float (&v)[4])
{
    const int g = lane >> 4;
    if constexpr (ROWC) {
        const float* f = reinterpret_cast<const float*>(S);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int kr = kk + 4 * g + s;
            v[s] = f[kr * RT + ((x + rd_rot(kr)) & (RT - 1))];
        }
    } else {
        const f32x4 t = *reinterpret_cast<const f32x4*>(S + x * (RK * 4) + (((kk / 4 + g) ^ ((x >> 1) & 7)) << 4));
#pragma unroll
        for (int s = 0; s < 4; ++s) v[s] = t[s];
    }
}

template <bool PT, bool QT>
__global__ void __launch_bounds__(256, 2) residual_f64_dma_kernel(icl_residual_args a)
{
    extern __shared__ __attribute__((aligned(1024))) unsigned char rd_lds[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tiles_m = a.M / RT, tiles_n = a.N / RT;
    const int bid = blockIdx.x;
    int tm[2], tn[2], ntiles = 1;
    if (a.tri_on == 1) {
        tn[0] = tn[1] = bid % tiles_n;
        tm[0] = bid / tiles_n;
        tm[1] = tiles_m - 1 - tm[0];
        ntiles = tm[1] > tm[0] ? 2 : 1;
    } else if (a.tri_on == 2) {
        tm[0] = tm[1] = bid % tiles_m;
        tn[0] = bid / tiles_m;
        tn[1] = tiles_n - 1 - tn[0];
        ntiles = tn[1] > tn[0] ? 2 : 1;
    } else {
        tm[0] = tm[1] = bid % tiles_m;
        tn[0] = tn[1] = bid / tiles_m;
    }
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    // P(i, k): P[i * ldp + k] when PT (k contiguous), else P[i + k * ldp] (rows
    // contiguous); Q(k, c): Q[k * ldq + c] when QT (rows contiguous), else Q[c * ldq + k].
    const long p_si = PT ? a.ldp : 1, p_sk = PT ? 1 : a.ldp;
    const long q_si = QT ? 1 : a.ldq, q_sk = QT ? a.ldq : 1;

    for (int t = 0; t < ntiles; ++t) {
        const int m0 = tm[t] * RT, n0 = tn[t] * RT;
        int k_lo = 0, k_hi = a.K;
        if (a.tri_on == 1) {
            if (a.tri_lower) k_hi = min(a.K, m0 + RT);
            else k_lo = m0;
        } else if (a.tri_on == 2) {
            if (a.tri_lower) k_lo = n0;
            else k_hi = min(a.K, n0 + RT);
        }
        if (t > 0) __syncthreads(); // every wave is done with the previous tile's last slab

        f64x4 acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
        const int m_base = m0 + wm + (lane & 15), n_base = n0 + wn + (lane >> 4);
        float b0[2][2][4];

        auto issue = [&](int kb, int st) {
            unsigned char* S = rd_lds + st * kRdStage;
            rd_issue<!PT>(a.P, p_si, p_sk, m0, kb, S, wave, lane);
            rd_issue<QT>(a.Q, q_si, q_sk, n0, kb, S + kRdOpB, wave, lane);
        };
        int cur = 0;
        if (k_lo < k_hi) issue(k_lo, 0);
        for (int kb = k_lo; kb < k_hi; kb += RK) {
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            const bool more = kb + RK < k_hi;
            if (more) issue(kb + RK, cur ^ 1);
            else {
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            b0[i][j][r] = a.B0[(long)(n_base + 16 * j + 4 * r) * a.ldb0 + m_base + 16 * i];
            }
            const unsigned char* Ps = rd_lds + cur * kRdStage;
            const unsigned char* Qs = Ps + kRdOpB;
            // triangle rule only on a slab that crosses the diagonal (uniform branch)
            const bool p_diag = a.tri_on == 1 && kb < m0 + RT && m0 < kb + RK;
            const bool q_diag = a.tri_on == 2 && kb < n0 + RT && n0 < kb + RK;
#pragma unroll
            for (int kk = 0; kk < RK; kk += 16) {
                float pf[2][4], qf[2][4];
#pragma unroll
                for (int i = 0; i < 2; ++i) rd_frag<!PT>(Ps, wm + 16 * i + (lane & 15), kk, lane, pf[i]);
#pragma unroll
                for (int j = 0; j < 2; ++j) rd_frag<QT>(Qs, wn + 16 * j + (lane & 15), kk, lane, qf[j]);
                if (p_diag || q_diag) {
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int k = kb + kk + 4 * (lane >> 4) + s;
#pragma unroll
                        for (int i = 0; i < 2; ++i) {
                            const int x = m0 + wm + 16 * i + (lane & 15); // P(x, k)
                            if (p_diag) {
                                float v = (a.tri_lower ? k > x : k < x) ? 0.0f : pf[i][s];
                                pf[i][s] = (a.tri_unit && k == x) ? 1.0f : v;
                            }
                        }
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            const int c = n0 + wn + 16 * j + (lane & 15); // Q(k, c)
                            if (q_diag) {
                                float v = (a.tri_lower ? c > k : c < k) ? 0.0f : qf[j][s];
                                qf[j][s] = (a.tri_unit && k == c) ? 1.0f : v;
                            }
                        }
                    }
                }
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    double p[2], q[2];
#pragma unroll
                    for (int i = 0; i < 2; ++i) p[i] = (double)pf[i][s];
#pragma unroll
                    for (int j = 0; j < 2; ++j) q[j] = (double)qf[j][s];
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j) acc[i][j] = icl::mfma_16x16x4_f64(q[j], p[i], acc[i][j]);
                }
            }
            cur ^= 1;
        }
        if (k_lo >= k_hi) { // no slab: R = alpha * B0
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        b0[i][j][r] = a.B0[(long)(n_base + 16 * j + 4 * r) * a.ldb0 + m_base + 16 * i];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    a.R[(long)(n_base + 16 * j + 4 * r) * a.ldr + m_base + 16 * i] =
                        (float)(a.alpha * (double)b0[i][j][r] - acc[i][j][r]);
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
            if (g == 0 && r0 + row < a.nsys && col < ncols) a.X[(long)(r0 + row) * a.xv_si + (long)(c0 + col) * a.xv_sj] = x[r];
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

// ---------------------------------------------------------------------------
// Dataflow blocked solve: the whole solve in one launch, one workgroup per
// 64 x 64 tile (row block b, line block c) of X:
//   X_bc = inv(L_bb) (B_bc - sum_{j<b} L_bj X_jc)
// Tiles are taken in row-block order from a ticket counter, so a workgroup only
// waits for tiles whose tickets are smaller; those were taken by workgroups
// that are already running and themselves wait only on smaller tickets, so
// every wait ends (by induction from ticket 0) however many workgroups are
// resident. Each wait is also bounded: past a.wait_polls polls (kDfSpin by
// default) it sets the abort word, and every later wait returns at once, so the
// launch always drains.
// A wave polls up to 64 flags at a time and skips the polls (and acquires) of
// the steps whose producers it has already seen finished.
// A producer publishes a tile from one wave, with a release fence at agent
// scope between the tile's stores and its flag; consumers acquire at agent
// scope after seeing the flag (the tiles may be produced on another XCD, whose
// L2 is not coherent with this one).
// X is tile-major; a tile stores its 64 x 64 values with the dimension that is
// contiguous in B fastest (lines when LINES_CONTIG, rows otherwise).
// ---------------------------------------------------------------------------
constexpr int kDfLd = kNB + 1;                                      // pitch of the [k][r] LDS tiles
constexpr int kDfSpin = 1 << 16;                                    // default polls per wait before aborting
constexpr size_t kDfLds = sizeof(float) * ((size_t)2 * kNB * kDfLd + (size_t)kNB * kLdD);

__device__ __forceinline__ int df_poll(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every lane of a wave calls this with the same j and n. Waits until the flag of
// tile j (flags[j * stride]) is set, polling the flags of tiles j .. j + 63 (< n)
// at once, one per lane, and returns the end of the run of set flags starting at
// j, so the steps up to it need no further poll. One acquire at agent scope
// covers the whole run. Returns n when the solve was aborted (then nothing may be
// assumed about the tiles). At most max_polls polls: the bound is tested before
// each poll, so max_polls == 0 aborts the wait at once.
__device__ __forceinline__ int df_ready(const int* flags, long stride, int j, int n, int lane, int* abort_word,
                                        int* error_word, int max_polls)
{
    const int jj = j + lane;
    for (int it = 0;; ++it) {
        if (it >= max_polls) {
            __hip_atomic_store(abort_word, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(error_word, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return n;
        }
        const int f = jj < n ? df_poll(flags + (long)jj * stride) : 0;
        const unsigned long long set = __ballot(f != 0);
        const int run = ~set == 0ull ? 64 : __builtin_ctzll(~set);
        if (run > 0) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return j + run;
        }
        if (df_poll(abort_word) != 0) return n;
        __builtin_amdgcn_s_sleep(2);
    }
}

template <bool LINES_CONTIG>
__device__ __forceinline__ f32x16 df_mfma(float row_frag, float line_frag, f32x16 acc)
{
    // lanes of the result run along lines when LINES_CONTIG, along rows otherwise
    return LINES_CONTIG ? icl::mfma_32x32x2_f32(row_frag, line_frag, acc)
                        : icl::mfma_32x32x2_f32(line_frag, row_frag, acc);
}

template <bool LINES_CONTIG, bool L_ROWS_CONTIG>
__global__ void __launch_bounds__(256) trsm_dataflow_kernel(icl_trsm_blocked_args a, float alpha)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    float* Ds = sm;                  // Ds[k][r] = inv(L_bb)(r, k), pitch kLdD
    float* Ls = Ds + kNB * kLdD;     // Ls[k][r] = L(64 b + r, 64 j + k)
    float* Xs = Ls + kNB * kDfLd;    // Xs[k][l] = X_jc(k, l); at the end R(k, l)
    __shared__ int s_ticket;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = (a.nsys + kNB - 1) / kNB, ncb = (a.nlines + kNB - 1) / kNB;
    int* const flags = a.sync;
    int* const ticket = a.sync + nblk * ncb;
    int* const abort_word = ticket + 1;
    const int max_polls = a.wait_polls == 0 ? kDfSpin : (a.wait_polls < 0 ? 0 : a.wait_polls);
    if (tid == 0) s_ticket = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int t = s_ticket;
    if (t >= nblk * ncb) return; // not reached: the grid has exactly one workgroup per tile
    const int b = t / ncb, c = t % ncb;
    const int r0 = b * kNB, c0 = c * kNB;
    const int ncols = min(kNB, a.nlines - c0);
    // tile-major X: element (k, l) of a tile at k * s_k + l * s_l
    constexpr long s_l = LINES_CONTIG ? 1 : kNB, s_k = LINES_CONTIG ? kNB : 1;
    auto xtile = [&](int j) { return a.X + ((long)j * ncb + c) * (kNB * kNB); };

    // Static operands first: inv(L_bb) into LDS and this tile's B values in the
    // accumulator layout (no other tile reads or writes B_bc).
    const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1); // quadrant of this wave: rows wr.., lines wc..
    const int row_t = wr + (LINES_CONTIG ? 4 * (lane >> 5) : (lane & 31));
    const int col_t = wc + (LINES_CONTIG ? (lane & 31) : 4 * (lane >> 5));
    const float* b_org = a.B + (long)r0 * a.x_si + (long)c0 * a.x_sj;
    float* const pb = a.B + (long)(r0 + row_t) * a.x_si + (long)(c0 + col_t) * a.x_sj;
    float bv[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int d = 8 * (r >> 2) + (r & 3);
        const int dr = LINES_CONTIG ? d : 0, dc = LINES_CONTIG ? 0 : d;
        const bool ok = r0 + row_t + dr < a.nsys && col_t + dc < ncols;
        const float x = *(ok ? pb + dr * a.x_si + dc * a.x_sj : b_org);
        bv[r] = ok ? x : 0.0f;
    }
    {
        const float4* inv4 = reinterpret_cast<const float4*>(a.invD + (long)b * kNB * kNB);
#pragma unroll
        for (int i = 0; i < kNB * kNB / 4 / 256; ++i) {
            const int e = 4 * (tid + 256 * i);
            *reinterpret_cast<float4*>(Ds + (e / kNB) * kLdD + (e % kNB)) = inv4[tid + 256 * i];
        }
    }

    // sum_{j<b} L_bj X_jc. The (L, X) pair of step j + 1 is read into registers
    // during the multiplication of step j when its producer is already known to
    // have finished; otherwise step j is multiplied first and the wait for tile
    // j + 1 comes after it, so that only the read and the multiplication of the
    // newest tile lie between its publication and this tile's.
    const float* l_row = a.L + (long)r0 * a.l_si;
    const float* l_diag = l_row + (long)r0 * a.l_sk;
    TileRegs<kNB, kNB, L_ROWS_CONTIG> lt;
    TileRegs<kNB, kNB, LINES_CONTIG> xt;
    f32x16 acc = f32x16{};
    int ready = 0; // tiles (j, c) with j < ready are known to be published (uniform per wave)
    auto fetch = [&](int j) {
        lt.load(l_row + (long)j * kNB * a.l_sk, a.l_si, a.l_sk, a.nsys - r0, kNB, tid, l_diag);
        if (j >= ready) ready = df_ready(flags + c, ncb, j, b, lane, abort_word, a.error_word, max_polls);
        xt.load(xtile(j), s_l, s_k, kNB, kNB, tid, a.X);
    };
    auto multiply = [&]() {
#pragma unroll 8
        for (int kk = 0; kk < kNB; kk += 2) {
            const int kr = kk + (lane >> 5);
            acc = df_mfma<LINES_CONTIG>(Ls[kr * kDfLd + wr + (lane & 31)], Xs[kr * kDfLd + wc + (lane & 31)], acc);
        }
    };
    if (b > 0) fetch(0);
    for (int j = 0; j < b; ++j) {
        __syncthreads(); // every wave is done with the tiles of step j - 1
        lt.template store<kDfLd>(Ls, tid);
        xt.template store<kDfLd>(Xs, tid);
        __syncthreads();
        if (j + 1 < b && j + 1 < ready) { // producer done: overlap the read with this step
            fetch(j + 1);
            multiply();
        } else {
            multiply();
            if (j + 1 < b) fetch(j + 1);
        }
    }

    // R = B_bc - sum into Xs as R[k = row][line]; X_bc = inv(L_bb) R
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int d = 8 * (r >> 2) + (r & 3);
        const int row = row_t + (LINES_CONTIG ? d : 0), col = col_t + (LINES_CONTIG ? 0 : d);
        Xs[row * kDfLd + col] = bv[r] - acc[r];
    }
    __syncthreads();
    f32x16 x = f32x16{};
#pragma unroll 8
    for (int kk = 0; kk < kNB; kk += 2) {
        const int kr = kk + (lane >> 5);
        x = df_mfma<LINES_CONTIG>(Ds[kr * kLdD + wr + (lane & 31)], Xs[kr * kDfLd + wc + (lane & 31)], x);
    }

    // Publish X_bc with a single writer: the waves stage their quadrants in Ls
    // as X[row][line]; wave 0 stores the whole tile (rows / lines beyond the
    // matrix are zero), then its release fence at agent scope orders those stores
    // before the flag. Then every wave writes B = alpha X.
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int d = 8 * (r >> 2) + (r & 3);
        const int row = row_t + (LINES_CONTIG ? d : 0), col = col_t + (LINES_CONTIG ? 0 : d);
        Ls[row * kDfLd + col] = x[r];
    }
    __syncthreads();
    if (wave == 0) {
        float4* xo = reinterpret_cast<float4*>(xtile(b));
#pragma unroll
        for (int i = 0; i < kNB * kNB / 4 / 64; ++i) {
            const int q = lane + 64 * i, e = 4 * q;
            // tile element e..e+3 along the fast dimension: (k, l..l+3) or (k..k+3, l)
            const int fast = e % kNB, slow = e / kNB;
            float4 v;
            if (LINES_CONTIG) { // e = k * 64 + l
                v = make_float4(Ls[slow * kDfLd + fast], Ls[slow * kDfLd + fast + 1], Ls[slow * kDfLd + fast + 2],
                                Ls[slow * kDfLd + fast + 3]);
            } else { // e = l * 64 + k
                v = make_float4(Ls[fast * kDfLd + slow], Ls[(fast + 1) * kDfLd + slow], Ls[(fast + 2) * kDfLd + slow],
                                Ls[(fast + 3) * kDfLd + slow]);
            }
            xo[q] = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane == 0) __hip_atomic_store(&flags[t], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int d = 8 * (r >> 2) + (r & 3);
        const int dr = LINES_CONTIG ? d : 0, dc = LINES_CONTIG ? 0 : d;
        if (r0 + row_t + dr < a.nsys && col_t + dc < ncols)
            __builtin_nontemporal_store(alpha * x[r], pb + dr * a.x_si + dc * a.x_sj);
    }
}

// ---------------------------------------------------------------------------
// Panel solve: rows [ps, pe) of the canonical system (at most four 64-row
// blocks), one workgroup per 64 lines, entirely inside the workgroup: the
// panel of B is staged in LDS as P[row][line]; block i is solved as
// X_i = inv(L_ii) B_i and subtracted from the later blocks of the panel
// (B_k -= L_ki X_i), each a 64 x 64 x 64 product on v_mfma_f32_32x32x2_f32
// (one 32 x 32 quadrant per wave).
// ---------------------------------------------------------------------------
constexpr int kPanelRows = 4 * kNB;
constexpr int kPanelLd = kStepCols + 1;
constexpr size_t kPanelLds = sizeof(float) * ((size_t)kPanelRows * kPanelLd + (size_t)kNB * kDfLd);

template <bool LINES_CONTIG, bool L_ROWS_CONTIG>
__global__ void __launch_bounds__(256) trsm_panel_kernel(icl_trsm_blocked_args a, int ps, int pe)
{
    extern __shared__ __attribute__((aligned(16))) float sm[];
    float* P = sm;                      // P[row][line]: panel rows of B, then of X
    float* T = P + kPanelRows * kPanelLd; // T[k][r]: operand tile of the current product
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = blockIdx.x * kStepCols;
    const int ncols = min(kStepCols, a.nlines - c0);
    const int np = pe - ps;
    const int nsub = (np + kNB - 1) / kNB;
    const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1); // quadrant of this wave: rows wr.., lines wc..

    // the panel of B: element (line l, row k) at B[(ps + k) * x_si + (c0 + l) * x_sj]
    const float* bsrc = a.B + (long)ps * a.x_si + (long)c0 * a.x_sj;
    TileRegs<kNB, kStepCols, LINES_CONTIG> bt[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < nsub) bt[i].load(bsrc + (long)(kNB * i) * a.x_si, a.x_sj, a.x_si, ncols, np - kNB * i, tid, bsrc);

    // The operand tiles of the (at most ten) products, in product order
    // inv(L_00), L_10, L_20, L_30, inv(L_11), L_21, L_31, inv(L_22), L_32,
    // inv(L_33), are all read into registers up front, so the products wait for
    // one memory latency in total rather than one each. Product q works on
    // block kPi[q] of X and block kPk[q] of B (the inverse when they are equal).
    // L_ki is read through the L view, inv(L_ii) from its dense column-major
    // block (element (r, k) at invD[k * 64 + r]); either way the store gives
    // T[k][r] = operand(r, k).
    constexpr int kPi[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
    constexpr int kPk[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
    TileRegs<kNB, kNB, L_ROWS_CONTIG> ot[10];
    const float* l_base = a.L + (long)ps * (a.l_si + a.l_sk);
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        const int i = kPi[q], k = kPk[q];
        if (k < nsub) {
            if (i == k)
                ot[q].load(a.invD + (long)(ps / kNB + i) * kNB * kNB, 1, kNB, kNB, kNB, tid, a.invD);
            else
                ot[q].load(l_base + (long)(kNB * k) * a.l_si + (long)(kNB * i) * a.l_sk, a.l_si, a.l_sk,
                           np - kNB * k, kNB, tid, l_base);
        }
    }
    // with every read in flight, stage the panel of B
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i < nsub) bt[i].template store<kPanelLd>(P + kNB * i * kPanelLd, tid);
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        const int i = kPi[q], k = kPk[q];
        if (k >= nsub) continue; // uniform: the panel has fewer blocks
        __syncthreads();         // the previous product is done with T (and P is staged)
        ot[q].template store<kDfLd>(T, tid);
        __syncthreads();
        float* Xi = P + kNB * i * kPanelLd; // rows of block i
        if (k == i) {
            // X_i = inv(L_ii) B_i: lanes along rows of the result
            f32x16 x = f32x16{};
#pragma unroll 8
            for (int kk = 0; kk < kNB; kk += 2) {
                const int kr = kk + (lane >> 5);
                x = icl::mfma_32x32x2_f32(Xi[kr * kPanelLd + wc + (lane & 31)], T[kr * kDfLd + wr + (lane & 31)], x);
            }
            __syncthreads(); // every wave has read B_i
#pragma unroll
            for (int r = 0; r < 16; ++r)
                Xi[(wr + (lane & 31)) * kPanelLd + wc + icl::mfma32_row(lane, r)] = x[r];
        } else {
            // B_k -= L_ki X_i
            float* Bk = P + kNB * k * kPanelLd;
            f32x16 u;
#pragma unroll
            for (int r = 0; r < 16; ++r) u[r] = Bk[(wr + (lane & 31)) * kPanelLd + wc + icl::mfma32_row(lane, r)];
#pragma unroll 8
            for (int kk = 0; kk < kNB; kk += 2) {
                const int kr = kk + (lane >> 5);
                u = icl::mfma_32x32x2_f32(Xi[kr * kPanelLd + wc + (lane & 31)], -T[kr * kDfLd + wr + (lane & 31)], u);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) Bk[(wr + (lane & 31)) * kPanelLd + wc + icl::mfma32_row(lane, r)] = u[r];
        }
    }
    __syncthreads();

    // X rows [ps, pe) through the X view, lanes along its contiguous dimension:
    // 64 elements per thread, addresses from one base plus uniform strides.
    {
        float* dst = a.X + (long)ps * a.xv_si + (long)c0 * a.xv_sj;
        if (LINES_CONTIG) { // thread: line tid % 64, rows tid / 64 + 4 j
            const int l = tid % kStepCols, r0t = tid / kStepCols;
            float* d = dst + (long)r0t * a.xv_si + (long)l * a.xv_sj;
#pragma unroll
            for (int j = 0; j < kPanelRows / 4; ++j) {
                const int r = r0t + 4 * j;
                if (r < np && l < ncols) d[(long)(4 * j) * a.xv_si] = P[r * kPanelLd + l];
            }
        } else { // thread: row tid, all lines
            const int r = tid;
            float* d = dst + (long)r * a.xv_si;
#pragma unroll
            for (int l = 0; l < kStepCols; ++l)
                if (r < np && l < ncols) d[(long)l * a.xv_sj] = P[r * kPanelLd + l];
        }
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
        a.B[i * a.x_si + j * a.x_sj] = alpha * a.X[i * a.xv_si + j * a.xv_sj];
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
    hipLaunchKernelGGL(trsm_invert_kernel, dim3(blocks), dim3(256), 0, stream, *a);
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

extern "C" int icl_trsm_dataflow_sync_ints(int nsys, int nlines)
{
    return ((nsys + kNB - 1) / kNB) * ((nlines + kNB - 1) / kNB) + 2;
}

extern "C" hipError_t icl_trsm_dataflow(const icl_trsm_blocked_args* a, float alpha, hipStream_t stream)
{
    if (a->nsys <= 0 || a->nlines <= 0) return hipSuccess;
    if (!a->sync || !a->error_word) return hipErrorInvalidValue;
    const long tiles = (long)((a->nsys + kNB - 1) / kNB) * ((a->nlines + kNB - 1) / kNB);
    using Kern = void (*)(icl_trsm_blocked_args, float);
    static const Kern kernels[4] = {trsm_dataflow_kernel<false, false>, trsm_dataflow_kernel<false, true>,
                                    trsm_dataflow_kernel<true, false>, trsm_dataflow_kernel<true, true>};
    const bool lines_contig = a->x_sj == 1 || a->x_sj == -1;
    const bool l_rows_contig = a->l_si == 1 || a->l_si == -1;
    hipLaunchKernelGGL(kernels[2 * lines_contig + l_rows_contig], dim3((unsigned)tiles), dim3(256), kDfLds, stream, *a,
                       alpha);
    return hipGetLastError();
}

namespace {

using PanelKern = void (*)(icl_trsm_blocked_args, int, int);

hipError_t launch_panel(const icl_trsm_blocked_args* a, int ps, int pe, hipStream_t stream)
{
    using Kern = PanelKern;
    static const Kern kernels[4] = {trsm_panel_kernel<false, false>, trsm_panel_kernel<false, true>,
                                    trsm_panel_kernel<true, false>, trsm_panel_kernel<true, true>};
    static const hipError_t attr = [] {
        for (Kern k : kernels) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPanelLds);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }();
    if (attr != hipSuccess) return attr;
    const bool lines_contig = a->x_sj == 1 || a->x_sj == -1;
    const bool l_rows_contig = a->l_si == 1 || a->l_si == -1;
    const unsigned blocks = (unsigned)((a->nlines + kStepCols - 1) / kStepCols);
    hipLaunchKernelGGL(kernels[2 * lines_contig + l_rows_contig], dim3(blocks), dim3(256), kPanelLds, stream, *a, ps,
                       pe);
    return hipGetLastError();
}

} // namespace

extern "C" hipError_t icl_trsm_panel(const icl_trsm_blocked_args* a, int ps, int pe, hipStream_t stream)
{
    if (a->nlines <= 0 || pe <= ps) return hipSuccess;
    if (ps % kNB != 0 || pe - ps > kPanelRows || pe > a->nsys) return hipErrorInvalidValue;
    return launch_panel(a, ps, pe, stream);
}

namespace {

// Inverses of the full 256 x 256 diagonal panels by recursive doubling over
// their 64 x 64 blocks, from the inverted diagonal blocks iD_i: a lower
// triangular [A 0; C D] has the inverse [inv(A) 0; -inv(D) C inv(A) inv(D)].
// Three launches: level 0, workgroup (p, h), the 128 x 128 halves {2h, 2h + 1}
// (W_aa = iD_a, W_dd = iD_d, W_ad = 0, W_da = -iD_d (L_da iD_a); h = 0 also zeroes
// the upper right quadrant); level 1, workgroup (p, q), T_rc = sum_{k = c..1}
// L_rk W_kc (r = 2 + q / 2, c = q % 2); level 2, W_rc = -sum_{k = 2..r} W_rk T_kc.
// Each 64^3 product runs on v_mfma_f32_32x32x2_f32 with A staged as [k][row] and
// B as [k][col] in LDS, each of the four waves owning a 32 x 32 quadrant.
constexpr int kTpLd = kNB + 1;
constexpr size_t kTpLds = sizeof(float) * 2 * kNB * kTpLd;
constexpr int kTpPer = kNB * kNB / 256; // block elements per thread: e = tid + 256 q, (x = e % 64, y = e / 64)

struct TpBlk {
    enum Kind { L, INVD, WBLK, TBLK } kind;
    int bi, bj;
};

__global__ void __launch_bounds__(256) trsm_invert_panels_kernel(icl_trsm_panel_inv_args a, int level)
{
    __shared__ float As[kNB][kTpLd]; // A[k][row]
    __shared__ float Bs[kNB][kTpLd]; // B[k][col]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = 32 * (wave & 1), wc = 32 * (wave >> 1);
    const int pnl = blockIdx.x, ps = pnl * 256;
    float* W = a.W + (long)pnl * 65536;
    auto w_at = [&](int r, int c) -> long {
        return a.rev ? (long)(255 - r) * a.w_rs + (long)(255 - c) * a.w_cs : (long)r * a.w_rs + (long)c * a.w_cs;
    };
    auto t_blk = [&](int r, int c) { return a.T + ((long)pnl * 4 + (r - 2) * 2 + c) * kNB * kNB; };
    auto load = [&](const TpBlk& b, float (&v)[kTpPer]) {
#pragma unroll
        for (int q = 0; q < kTpPer; ++q) {
            const int e = tid + 256 * q, x = e % kNB, y = e / kNB;
            const int r = kNB * b.bi + x, c = kNB * b.bj + y;
            switch (b.kind) {
            case TpBlk::L: v[q] = a.L[(long)(ps + r) * a.l_si + (long)(ps + c) * a.l_sk]; break;
            case TpBlk::INVD: v[q] = a.invD[((long)(ps / kNB + b.bi) * kNB + y) * kNB + x]; break;
            case TpBlk::WBLK: v[q] = W[w_at(r, c)]; break;
            default: v[q] = t_blk(b.bi, b.bj)[(long)y * kNB + x]; break;
            }
        }
    };
    // acc += X Y with X staged as the A tile and Y as the B tile (after a
    // barrier, so that no wave still reads the previous tiles).
    auto product = [&](const TpBlk& X, const TpBlk& Y, f32x16& acc) {
        float va[kTpPer], vb[kTpPer];
        load(X, va);
        load(Y, vb);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kTpPer; ++q) {
            const int e = tid + 256 * q, x = e % kNB, y = e / kNB; // X: (row x, k y); Y: (k x, col y)
            As[y][x] = va[q];
            Bs[x][y] = vb[q];
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < kNB; kk += 2) {
            const int kr = kk + (lane >> 5);
            acc = icl::mfma_32x32x2_f32(Bs[kr][wc + (lane & 31)], As[kr][wr + (lane & 31)], acc);
        }
    };
    auto row_of = [&]() { return wr + (lane & 31); };
    auto col_of = [&](int r) { return wc + icl::mfma32_row(lane, r); };
    auto store_w = [&](int bi, int bj, const f32x16& acc) { // W_(bi, bj) = -acc
#pragma unroll
        for (int r = 0; r < 16; ++r) W[w_at(kNB * bi + row_of(), kNB * bj + col_of(r))] = -acc[r];
    };
    auto copy_w = [&](bool inv, int bi, int bj) { // W_(bi, bj) = iD_bi (inv) or 0
        float v[kTpPer];
        if (inv) load({TpBlk::INVD, bi, bi}, v);
        else
            for (int q = 0; q < kTpPer; ++q) v[q] = 0.0f;
#pragma unroll
        for (int q = 0; q < kTpPer; ++q) {
            const int e = tid + 256 * q;
            W[w_at(kNB * bi + e % kNB, kNB * bj + e / kNB)] = v[q];
        }
    };

    if (level == 0) {
        const int h = blockIdx.y, da = 2 * h, dd = da + 1;
        copy_w(true, da, da);
        copy_w(true, dd, dd);
        copy_w(false, da, dd);
        if (h == 0)
            for (int i = 0; i < 2; ++i)
                for (int j = 2; j < 4; ++j) copy_w(false, i, j);
        f32x16 s = f32x16{};
        product({TpBlk::L, dd, da}, {TpBlk::INVD, da, da}, s); // S = L_da iD_a
        // W_da = -iD_d S: S becomes the B tile, iD_d the A tile.
        float vd[kTpPer];
        load({TpBlk::INVD, dd, dd}, vd);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; ++r) Bs[row_of()][col_of(r)] = s[r];
#pragma unroll
        for (int q = 0; q < kTpPer; ++q) {
            const int e = tid + 256 * q;
            As[e / kNB][e % kNB] = vd[q];
        }
        __syncthreads();
        f32x16 m = f32x16{};
#pragma unroll 8
        for (int kk = 0; kk < kNB; kk += 2) {
            const int kr = kk + (lane >> 5);
            m = icl::mfma_32x32x2_f32(Bs[kr][wc + (lane & 31)], As[kr][wr + (lane & 31)], m);
        }
        store_w(dd, da, m);
        return;
    }
    const int q = blockIdx.y, r = 2 + (q >> 1), c = q & 1;
    f32x16 acc = f32x16{};
    if (level == 1) { // T_rc = sum_{k = c..1} L_rk W_kc
        for (int k = c; k < 2; ++k) product({TpBlk::L, r, k}, {TpBlk::WBLK, k, c}, acc);
        float* t = t_blk(r, c);
#pragma unroll
        for (int e = 0; e < 16; ++e) t[(long)col_of(e) * kNB + row_of()] = acc[e];
        return;
    }
    for (int k = 2; k <= r; ++k) product({TpBlk::WBLK, r, k}, {TpBlk::TBLK, k, c}, acc);
    store_w(r, c, acc);
}

} // namespace

extern "C" hipError_t icl_trsm_invert_panels(const icl_trsm_panel_inv_args* a, hipStream_t stream)
{
    if (a->npanels <= 0) return hipSuccess;
    for (int level = 0; level < 3; ++level) {
        hipLaunchKernelGGL(trsm_invert_panels_kernel, dim3((unsigned)a->npanels, level == 0 ? 2 : 4), dim3(256), 0,
                           stream, *a, level);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
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
    const long tiles_m = (a->M + RT - 1) / RT, tiles_n = (a->N + RT - 1) / RT;
    // one workgroup per pair of tiles along the triangular operand's free dimension
    const long tiles = a->tri_on == 1 ? (tiles_m + 1) / 2 * tiles_n
                     : a->tri_on == 2 ? tiles_m * ((tiles_n + 1) / 2)
                                      : tiles_m * tiles_n;
    using Kern = void (*)(icl_residual_args);
    const int which = 2 * (a->p_trans != 0) + (a->q_trans != 0);
    // The DMA kernel: no edges, and 16-byte aligned operands whose strides are
    // multiples of 4 floats.
    auto al16 = [](const void* p, long ld) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0 && ld % 4 == 0; };
    if (a->M % RT == 0 && a->N % RT == 0 && a->K % RK == 0 && al16(a->P, a->ldp) && al16(a->Q, a->ldq)) {
        static const Kern dma[4] = {residual_f64_dma_kernel<false, false>, residual_f64_dma_kernel<false, true>,
                                    residual_f64_dma_kernel<true, false>, residual_f64_dma_kernel<true, true>};
        hipLaunchKernelGGL(dma[which], dim3((unsigned)tiles), dim3(256), kRdLds, stream, *a);
        return hipGetLastError();
    }
    static const Kern kernels[4] = {residual_f64_kernel<false, false>, residual_f64_kernel<false, true>,
                                    residual_f64_kernel<true, false>, residual_f64_kernel<true, true>};
    hipLaunchKernelGGL(kernels[which], dim3((unsigned)tiles), dim3(256), 0, stream, *a);
    return hipGetLastError();
}

namespace {

__global__ void __launch_bounds__(256) absmax_pair_kernel(int m, int n, const float* D, long ldd, const float* X,
                                                          long ldx, float* partial)
{
    __shared__ float sd[256], sx[256];
    // bpc blocks share a column (more than one when there are fewer columns than
    // blocks); block b takes part b % bpc of the rows of columns b / bpc,
    // b / bpc + cstep, ...
    const int bpc = n >= (int)gridDim.x ? 1 : (int)gridDim.x / n;
    const int cstep = (int)gridDim.x / bpc, part = blockIdx.x % bpc, c0 = blockIdx.x / bpc;
    float md = 0.0f, mx = 0.0f;
    if (c0 < cstep) {
        for (int j = c0; j < n; j += cstep) {
            const float* dc = D + (long)j * ldd;
            const float* xc = X + (long)j * ldx;
            const int step = bpc * 256;
            int i = part * 256 + (int)threadIdx.x;
            for (; i + 3 * step < m; i += 4 * step) { // four rows of each matrix in flight
                const float d0 = dc[i], d1 = dc[i + step], d2 = dc[i + 2 * step], d3 = dc[i + 3 * step];
                const float x0 = xc[i], x1 = xc[i + step], x2 = xc[i + 2 * step], x3 = xc[i + 3 * step];
                md = fmaxf(md, fmaxf(fmaxf(fabsf(d0), fabsf(d1)), fmaxf(fabsf(d2), fabsf(d3))));
                mx = fmaxf(mx, fmaxf(fmaxf(fabsf(x0), fabsf(x1)), fmaxf(fabsf(x2), fabsf(x3))));
            }
            for (; i < m; i += step) {
                md = fmaxf(md, fabsf(dc[i]));
                mx = fmaxf(mx, fabsf(xc[i]));
            }
        }
    }
    sd[threadIdx.x] = md;
    sx[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sd[threadIdx.x] = fmaxf(sd[threadIdx.x], sd[threadIdx.x + s]);
            sx[threadIdx.x] = fmaxf(sx[threadIdx.x], sx[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sd[0];
        partial[2 * blockIdx.x + 1] = sx[0];
    }
}

} // namespace

extern "C" hipError_t icl_absmax_pair(int m, int n, const float* D, long ldd, const float* X, long ldx, float* partial,
                                      hipStream_t stream)
{
    if (m <= 0 || n <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(absmax_pair_kernel, dim3(ICL_ABSMAX_BLOCKS), dim3(256), 0, stream, m, n, D, ldd, X, ldx,
                       partial);
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
