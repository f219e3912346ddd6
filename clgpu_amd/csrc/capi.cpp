// C entry points of the hot path: iclblasCreate/Destroy, iclblasSgemm,
// iclblasCgemm, iclblasStrsm, iclblasCtrsm, and the iclBLAS_ext.h extensions.
//
// Argument checks run in the reference's order and return before any HIP call
// (reference: libs/blas/shlib/iclBLAS_S.cpp:487-510, iclBLAS_C.cpp:458-469,
// iclBLASImpl.hpp:108-245, iclBLAS.cpp:35-59):
//   GEMM: m, n or k < 0 -> INVALID_VALUE; n == 0 or m == 0 -> SUCCESS (before the
//   handle is looked at); null A / lda <= 0 / null B / ldb <= 0 / null C /
//   ldc <= 0 -> INVALID_VALUE; invalid handle -> INVALID_VALUE.
//   TRSM: m == 0 or n == 0 -> SUCCESS first, then the same chain.
// Checks the reference does not make (it crashes or reads out of bounds instead)
// run after the reference's own ones and before the handle check, and return
// INVALID_VALUE: null alpha/beta, enum values out of range, lda/ldb/ldc smaller
// than the number of rows of the matrix they describe.
// No exception crosses this boundary (the reference's debug build lets them,
// iclBLASImpl.hpp:69-72).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "capi_common.hpp"

using icl::status_error;

using namespace icl::capi;

namespace {

// STRSM implementations above 64 unknowns (forceable by name): the panel solve
// with inverted 256 x 256 panels (default: each panel one GEMM with its inverse,
// the rest updated by a GEMM), the panel solve with the in-workgroup panel
// kernel, the launch-per-64-block solve with inverted diagonal blocks, the
// single-launch dataflow solve, and the recursive split into diagonal-block
// solves plus GEMM updates. CTRSM above 64 unknowns: the panel solve with
// inverted 256 x 256 panels (default) and the recursive split whose leaves are
// products with the inverted 64 x 64 diagonal blocks.
const char* const kTrsmImpls[] = {"strsm_blocked_inv64", "strsm_dataflow64",      "strsm_panel256",
                                  "strsm_panel256_inv",  "strsm_panel256_fused",  "strsm_recursive",
                                  "ctrsm_panel256_inv",  "ctrsm_recursive_inv64"};

// Shared GEMM validation; *work is set when the call has work to do.
iclblasStatus_t gemm_checks(iclblasHandle_t handle, int transa, int transb, int m, int n, int k, const void* alpha,
                            const void* A, int lda, const void* B, int ldb, const void* beta, const void* C, int ldc,
                            bool* work)
{
    *work = false;
    if (m < 0 || n < 0 || k < 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!alpha || !beta) return ICLBLAS_STATUS_INVALID_VALUE;
    if (n == 0 || m == 0) return ICLBLAS_STATUS_SUCCESS;
    if (!A || lda <= 0 || !B || ldb <= 0 || !C || ldc <= 0) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!op_ok(transa) || !op_ok(transb)) return ICLBLAS_STATUS_INVALID_VALUE;
    const int rows_a = transa == ICLBLAS_OP_N ? m : k;
    const int rows_b = transb == ICLBLAS_OP_N ? k : n;
    if (lda < (rows_a > 1 ? rows_a : 1) || ldb < (rows_b > 1 ? rows_b : 1) || ldc < m) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    *work = true;
    return ICLBLAS_STATUS_SUCCESS;
}

} // namespace

extern "C" {

ICLBLAS_API iclblasStatus_t iclblasCreate(iclblasHandle_t* handle)
{
    if (!handle) return ICLBLAS_STATUS_INVALID_VALUE;
    *handle = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return ICLBLAS_STATUS_NOT_INITIALIZED;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ICLBLAS_STATUS_NOT_INITIALIZED;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return ICLBLAS_STATUS_NOT_INITIALIZED;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return ICLBLAS_STATUS_ARCH_MISMATCH;
    iclblasContext* h = new (std::nothrow) iclblasContext();
    if (!h) return ICLBLAS_STATUS_ALLOC_FAILED;
    h->device = dev;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return ICLBLAS_STATUS_NOT_INITIALIZED;
    }
    h->own_stream = true;
    for (auto& s : h->scratch) s.set_budget(&h->budget);
    // The implementation name is copied on every call: room for the longest one
    // now, so that calls do not allocate host memory for it.
    h->last_impl.reserve(64);
    // iclblasSgemmNoAlloc's static scratch on this device (once per device; a
    // failure is reported by that entry point, not here).
    (void)icl_noalloc_scratch_init(dev);
    *handle = h;
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasDestroy(iclblasHandle_t handle)
{
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    (void)hipSetDevice(handle->device);
    (void)hipStreamSynchronize(handle->stream);
    handle->copy.release();
    handle->side.release();
    for (auto& s : handle->scratch) s.release();
    if (handle->trsm_error) (void)hipFree(handle->trsm_error);
    if (handle->own_stream) (void)hipStreamDestroy(handle->stream);
    handle->tag = 0;
    delete handle;
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasSgemm(iclblasHandle_t handle, iclblasOperation_t transa_, iclblasOperation_t transb_,
                                         int m, int n, int k, const float* alpha, float* A, int lda, float* B, int ldb,
                                         const float* beta, float* C, int ldc)
{
    const int transa = enum_arg(transa_);
    const int transb = enum_arg(transb_);
    bool work = false;
    const iclblasStatus_t st = gemm_checks(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, &work);
    if (!work) return st;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        const float a = *alpha, b = *beta;
        const int ra = transa == ICLBLAS_OP_N ? m : k, ca = transa == ICLBLAS_OP_N ? k : m;
        const int rb = transb == ICLBLAS_OP_N ? k : n, cb = transb == ICLBLAS_OP_N ? n : k;
        if (h->memory_mode == ICLBLAS_MEMORY_DEVICE) {
            icl::sgemm_dev(h, math_of(h), transa, transb, m, n, k, a, A, lda, B, ldb, b, C, ldc);
        } else if (const int width = host_gemm_panel_width(h, m, n, k)) {
            const bool use_ab = k > 0 && a != 0.0f;
            host_gemm_panels<float>(h, transb, m, n, k, use_ab, b != 0.0f, A, ra, ca, lda, B, rb, cb, ldb, C, ldc, width,
                                    [&](const float* dA, const float* dB, float* dC, int, int w) {
                                        icl::sgemm_dev(h, math_of(h), transa, transb, m, w, k, a, dA, lda, dB, ldb, b,
                                                       dC, ldc);
                                    });
        } else {
            const bool use_ab = k > 0 && a != 0.0f;
            const float* dA = use_ab ? stage_in<float>(h, kSlotA, A, ra, ca, lda, true) : nullptr;
            const float* dB = use_ab ? stage_in<float>(h, kSlotB, B, rb, cb, ldb, true) : nullptr;
            float* dC = stage_in<float>(h, kSlotC, C, m, n, ldc, b != 0.0f);
            icl::sgemm_dev(h, math_of(h), transa, transb, m, n, k, a, dA, lda, dB, ldb, b, dC, ldc);
            stage_out<float>(h, C, dC, m, n, ldc);
        }
        finish(h);
    });
}

ICLBLAS_API iclblasStatus_t iclblasSgemmNoAlloc(iclblasHandle_t handle, iclblasOperation_t transa_,
                                                iclblasOperation_t transb_, int m, int n, int k, const float* alpha,
                                                float* A, int lda, float* B, int ldb, const float* beta, float* C,
                                                int ldc)
{
    const int transa = enum_arg(transa_);
    const int transb = enum_arg(transb_);
    bool work = false;
    const iclblasStatus_t st = gemm_checks(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, &work);
    if (!work) return st;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        void* ws = icl_noalloc_scratch(h->device);
        if (!ws) throw status_error(ICLBLAS_STATUS_NOT_INITIALIZED, "no static scratch on this device");
        // The scratch replaces the handle's grow-only workspace for this call only.
        struct ExtWorkspace {
            iclblasContext* h;
            ExtWorkspace(iclblasContext* hh, void* p, size_t bytes) : h(hh)
            {
                h->ext_ws = p;
                h->ext_ws_bytes = bytes;
            }
            ~ExtWorkspace()
            {
                h->ext_ws = nullptr;
                h->ext_ws_bytes = 0;
            }
        } guard(h, ws, ICLBLAS_NOALLOC_SCRATCH_BYTES);
        icl::sgemm_dev(h, math_of(h), transa, transb, m, n, k, *alpha, A, lda, B, ldb, *beta, C, ldc);
        icl::hip_check(hipGetLastError(), "launch");
    });
}

ICLBLAS_API iclblasStatus_t iclblasCgemm(iclblasHandle_t handle, iclblasOperation_t transa_, iclblasOperation_t transb_,
                                         int m, int n, int k, const oclComplex_t* alpha, oclComplex_t* A, int lda,
                                         oclComplex_t* B, int ldb, const oclComplex_t* beta, oclComplex_t* C, int ldc)
{
    const int transa = enum_arg(transa_);
    const int transb = enum_arg(transb_);
    bool work = false;
    const iclblasStatus_t st = gemm_checks(handle, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, &work);
    if (!work) return st;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        float a[2], b[2];
        std::memcpy(a, alpha, sizeof(a));
        std::memcpy(b, beta, sizeof(b));
        typedef float cf[2];
        const int ra = transa == ICLBLAS_OP_N ? m : k, ca = transa == ICLBLAS_OP_N ? k : m;
        const int rb = transb == ICLBLAS_OP_N ? k : n, cb = transb == ICLBLAS_OP_N ? n : k;
        if (h->memory_mode == ICLBLAS_MEMORY_DEVICE) {
            icl::cgemm_dev(h, math_of(h), transa, transb, m, n, k, a, reinterpret_cast<const float*>(A), lda,
                           reinterpret_cast<const float*>(B), ldb, b, reinterpret_cast<float*>(C), ldc);
        } else if (const int width = host_gemm_panel_width(h, m, n, k)) {
            const bool use_ab = k > 0 && (a[0] != 0.0f || a[1] != 0.0f);
            const bool use_c = b[0] != 0.0f || b[1] != 0.0f;
            host_gemm_panels<cf>(h, transb, m, n, k, use_ab, use_c, reinterpret_cast<const cf*>(A), ra, ca, lda,
                                 reinterpret_cast<const cf*>(B), rb, cb, ldb, reinterpret_cast<cf*>(C), ldc, width,
                                 [&](const cf* dA, const cf* dB, cf* dC, int, int w) {
                                     icl::cgemm_dev(h, math_of(h), transa, transb, m, w, k, a,
                                                    reinterpret_cast<const float*>(dA), lda,
                                                    reinterpret_cast<const float*>(dB), ldb, b,
                                                    reinterpret_cast<float*>(dC), ldc);
                                 });
        } else {
            const bool use_ab = k > 0 && (a[0] != 0.0f || a[1] != 0.0f);
            const bool use_c = b[0] != 0.0f || b[1] != 0.0f;
            const cf* dA = use_ab ? stage_in<cf>(h, kSlotA, reinterpret_cast<const cf*>(A), ra, ca, lda, true) : nullptr;
            const cf* dB = use_ab ? stage_in<cf>(h, kSlotB, reinterpret_cast<const cf*>(B), rb, cb, ldb, true) : nullptr;
            cf* dC = stage_in<cf>(h, kSlotC, reinterpret_cast<const cf*>(C), m, n, ldc, use_c);
            icl::cgemm_dev(h, math_of(h), transa, transb, m, n, k, a, reinterpret_cast<const float*>(dA), lda,
                           reinterpret_cast<const float*>(dB), ldb, b, reinterpret_cast<float*>(dC), ldc);
            stage_out<cf>(h, reinterpret_cast<cf*>(C), dC, m, n, ldc);
        }
        finish(h);
    });
}

static iclblasStatus_t strsm_common(iclblasHandle_t handle, int side, int uplo, int trans, int diag, int m, int n,
                                    const float* alpha,
                                    float* A, int lda, float* B, int ldb, int sweeps)
{
    bool work = false;
    const iclblasStatus_t st = trsm_checks(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, &work);
    if (!work) return st;
    if (sweeps < 0 || sweeps > 8) return ICLBLAS_STATUS_INVALID_VALUE;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        h->last_sweeps = 0; // a call without sweeps (or with alpha == 0) ran none
        const float a = *alpha;
        const int order = side == ICLBLAS_SIDE_LEFT ? m : n;
        const float* dA = A;
        float* dB = B;
        const bool host = h->memory_mode == ICLBLAS_MEMORY_HOST;
        if (host) {
            dA = a != 0.0f ? stage_in<float>(h, kSlotA, A, order, order, lda, true) : nullptr;
            dB = stage_in<float>(h, kSlotB, B, m, n, ldb, a != 0.0f);
        }
        if (sweeps > 0 && a != 0.0f) {
            const size_t bytes = (size_t)m * n * sizeof(float);
            float* w0 = static_cast<float*>(h->scratch[kSlotW0].get(bytes, h->stream));
            float* w1 = static_cast<float*>(h->scratch[kSlotW1].get(bytes, h->stream));
            icl::strsm_refined_dev(h, side, uplo, trans, diag, m, n, a, dA, lda, dB, ldb, sweeps, w0, w1);
        } else {
            icl::strsm_dev(h, side, uplo, trans, diag, m, n, a, dA, lda, dB, ldb);
        }
        if (host) stage_out<float>(h, B, dB, m, n, ldb);
        finish(h);
        if (host) icl::trsm_check_error(h);
    });
}

// reference: libs/blas/shlib/iclBLAS_C.cpp:525-535 (m == 0 or n == 0 -> SUCCESS
// first, then the STRSM checks).
ICLBLAS_API iclblasStatus_t iclblasCtrsm(iclblasHandle_t handle, iclblasSideMode_t side_, iclblasFillMode_t uplo_,
                                         iclblasOperation_t trans_, iclblasDiagType_t diag_, int m, int n,
                                         const oclComplex_t* alpha, oclComplex_t* A, int lda, oclComplex_t* B, int ldb)
{
    const int trans = enum_arg(trans_);
    const int side = enum_arg(side_);
    const int uplo = enum_arg(uplo_);
    const int diag = enum_arg(diag_);
    bool work = false;
    const iclblasStatus_t st =
        trsm_checks(handle, side, uplo, trans, diag, m, n, reinterpret_cast<const float*>(alpha),
                    reinterpret_cast<const float*>(A), lda, reinterpret_cast<const float*>(B), ldb, &work);
    if (!work) return st;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        float a[2];
        std::memcpy(a, alpha, sizeof(a));
        typedef float cf[2];
        const int order = side == ICLBLAS_SIDE_LEFT ? m : n;
        const bool nonzero = a[0] != 0.0f || a[1] != 0.0f;
        const float* dA = reinterpret_cast<const float*>(A);
        float* dB = reinterpret_cast<float*>(B);
        const bool host = h->memory_mode == ICLBLAS_MEMORY_HOST;
        if (host) {
            dA = nonzero ? reinterpret_cast<const float*>(
                               stage_in<cf>(h, kSlotA, reinterpret_cast<const cf*>(A), order, order, lda, true))
                         : nullptr;
            dB = reinterpret_cast<float*>(stage_in<cf>(h, kSlotB, reinterpret_cast<const cf*>(B), m, n, ldb, nonzero));
        }
        icl::ctrsm_dev(h, side, uplo, trans, diag, m, n, a, dA, lda, dB, ldb);
        if (host) stage_out<cf>(h, reinterpret_cast<cf*>(B), reinterpret_cast<const cf*>(dB), m, n, ldb);
        finish(h);
    });
}

ICLBLAS_API iclblasStatus_t iclblasStrsm(iclblasHandle_t handle, iclblasSideMode_t side_, iclblasFillMode_t uplo_,
                                         iclblasOperation_t trans_, iclblasDiagType_t diag_, int m, int n,
                                         const float* alpha, float* A, int lda, float* B, int ldb)
{
    const int trans = enum_arg(trans_);
    const int side = enum_arg(side_);
    const int uplo = enum_arg(uplo_);
    const int diag = enum_arg(diag_);
    return strsm_common(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, 0);
}

ICLBLAS_API iclblasStatus_t iclblasStrsmRefined(iclblasHandle_t handle, iclblasSideMode_t side_, iclblasFillMode_t uplo_,
                                                iclblasOperation_t trans_, iclblasDiagType_t diag_, int m, int n,
                                                const float* alpha, float* A, int lda, float* B, int ldb,
                                                int max_sweeps)
{
    const int trans = enum_arg(trans_);
    const int side = enum_arg(side_);
    const int uplo = enum_arg(uplo_);
    const int diag = enum_arg(diag_);
    return strsm_common(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, max_sweeps);
}

// ---------------------------------------------------------------------------
// Modes, stream, implementation override
// ---------------------------------------------------------------------------
ICLBLAS_API iclblasStatus_t iclblasSetMemoryMode(iclblasHandle_t handle, iclblasMemoryMode_t mode_)
{
    const int mode = enum_arg(mode_);
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    if (mode != ICLBLAS_MEMORY_HOST && mode != ICLBLAS_MEMORY_DEVICE) return ICLBLAS_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> g(handle->mu);
    handle->memory_mode = static_cast<iclblasMemoryMode_t>(mode);
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasGetMemoryMode(iclblasHandle_t handle, iclblasMemoryMode_t* mode)
{
    if (!iclblasContext::valid(handle) || !mode) return ICLBLAS_STATUS_INVALID_VALUE;
    *mode = handle->memory_mode;
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasSetMathMode(iclblasHandle_t handle, iclblasMathMode_t mode_)
{
    const int mode = enum_arg(mode_);
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    if (mode != ICLBLAS_MATH_FP32 && mode != ICLBLAS_MATH_SPLIT_BF16X3) return ICLBLAS_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> g(handle->mu);
    handle->math_mode = static_cast<iclblasMathMode_t>(mode);
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasGetMathMode(iclblasHandle_t handle, iclblasMathMode_t* mode)
{
    if (!iclblasContext::valid(handle) || !mode) return ICLBLAS_STATUS_INVALID_VALUE;
    *mode = handle->math_mode;
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasSetStream(iclblasHandle_t handle, void* stream)
{
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> g(handle->mu);
    (void)hipSetDevice(handle->device);
    (void)hipStreamSynchronize(handle->stream);
    if (handle->own_stream) (void)hipStreamDestroy(handle->stream);
    handle->stream = static_cast<hipStream_t>(stream);
    handle->own_stream = false;
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasGetStream(iclblasHandle_t handle, void** stream)
{
    if (!iclblasContext::valid(handle) || !stream) return ICLBLAS_STATUS_INVALID_VALUE;
    *stream = handle->stream;
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasSynchronize(iclblasHandle_t handle)
{
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    return guarded([&] {
        HandleLock lock(handle);
        icl::hip_check(hipStreamSynchronize(handle->stream), "hipStreamSynchronize");
        icl::trsm_check_error(handle); // an aborted dataflow STRSM queued in the device-pointer mode
    });
}

ICLBLAS_API iclblasStatus_t iclblasSetImplementation(iclblasHandle_t handle, const char* name)
{
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> g(handle->mu);
    if (!name || !*name) {
        handle->forced_impl.clear();
        return ICLBLAS_STATUS_SUCCESS;
    }
    for (const icl_gemm_variant* v : icl::gemm_registry())
        if (std::strcmp(v->name, name) == 0) {
            handle->forced_impl = name;
            return ICLBLAS_STATUS_SUCCESS;
        }
    for (const char* t : kTrsmImpls)
        if (std::strcmp(t, name) == 0) {
            handle->forced_impl = name;
            return ICLBLAS_STATUS_SUCCESS;
        }
    return ICLBLAS_STATUS_INVALID_VALUE;
}

ICLBLAS_API iclblasStatus_t iclblasGetImplementations(char* buf, int len)
{
    if (!buf || len <= 0) return ICLBLAS_STATUS_INVALID_VALUE;
    std::string all;
    for (const icl_gemm_variant* v : icl::gemm_registry()) {
        if (!all.empty()) all += '\n';
        all += v->name;
    }
    for (const char* t : kTrsmImpls) {
        all += '\n';
        all += t;
    }
    if ((int)all.size() + 1 > len) return ICLBLAS_STATUS_INVALID_VALUE;
    std::memcpy(buf, all.c_str(), all.size() + 1);
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasGetLastImplementation(iclblasHandle_t handle, char* buf, int len)
{
    if (!iclblasContext::valid(handle) || !buf || len <= 0) return ICLBLAS_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> g(handle->mu);
    std::strncpy(buf, handle->last_impl.c_str(), (size_t)len - 1);
    buf[len - 1] = '\0';
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasGetLastRefinementSweeps(iclblasHandle_t handle, int* sweeps)
{
    if (!iclblasContext::valid(handle) || !sweeps) return ICLBLAS_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> g(handle->mu);
    *sweeps = handle->last_sweeps;
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasSetKernelTiming(iclblasHandle_t handle, int enable)
{
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    return guarded([&] {
        HandleLock lock(handle);
        icl::hip_check(hipStreamSynchronize(handle->stream), "hipStreamSynchronize");
        handle->timer.on = enable != 0;
        handle->timer.used = 0;
        handle->timer.dropped = 0;
    });
}

ICLBLAS_API iclblasStatus_t iclblasGetKernelTiming(iclblasHandle_t handle, float* total_ms, int* launches)
{
    if (!iclblasContext::valid(handle) || !total_ms || !launches) return ICLBLAS_STATUS_INVALID_VALUE;
    return guarded([&] {
        HandleLock lock(handle);
        icl::KernelTimer& t = handle->timer;
        icl::hip_check(hipStreamSynchronize(handle->stream), "hipStreamSynchronize");
        double sum = 0.0;
        for (size_t i = 0; i + 1 < t.used; i += 2) {
            icl::hip_check(hipEventSynchronize(t.ev[i + 1]), "hipEventSynchronize");
            float ms = 0.0f;
            icl::hip_check(hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]), "hipEventElapsedTime");
            sum += ms;
        }
        *total_ms = (float)sum;
        *launches = (int)(t.used / 2); // launches beyond the event cap are not in the sum
        t.used = 0;
        t.dropped = 0;
    });
}

ICLBLAS_API iclblasStatus_t iclblasSetScratchLimit(iclblasHandle_t handle, size_t bytes)
{
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> g(handle->mu);
    handle->budget.limit = bytes;
    return ICLBLAS_STATUS_SUCCESS;
}

ICLBLAS_API iclblasStatus_t iclblasGetScratchUsage(iclblasHandle_t handle, size_t* held, size_t* limit)
{
    if (!iclblasContext::valid(handle) || !held || !limit) return ICLBLAS_STATUS_INVALID_VALUE;
    std::lock_guard<std::mutex> g(handle->mu);
    *held = handle->budget.held;
    *limit = handle->budget.limit;
    return ICLBLAS_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// Adjoints
// ---------------------------------------------------------------------------
ICLBLAS_API iclblasStatus_t iclblasSgemmBackward(iclblasHandle_t handle, iclblasOperation_t transa_,
                                                 iclblasOperation_t transb_, int m, int n, int k, const float* alpha,
                                                 float* A, int lda, float* B, int ldb, float* dC, int lddc, float* dA,
                                                 int ldda, float* dB, int lddb)
{
    const int transa = enum_arg(transa_);
    const int transb = enum_arg(transb_);
    if (m < 0 || n < 0 || k < 0 || !alpha) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!op_ok(transa) || !op_ok(transb)) return ICLBLAS_STATUS_INVALID_VALUE;
    const int ra = transa == ICLBLAS_OP_N ? m : k, ca = transa == ICLBLAS_OP_N ? k : m;
    const int rb = transb == ICLBLAS_OP_N ? k : n, cb = transb == ICLBLAS_OP_N ? n : k;
    if (m == 0 || n == 0 || k == 0) {
        // Gradients of empty products: nothing flows (dA/dB would be zero-sized or
        // zero); keep the contract simple and require non-empty shapes.
        return ICLBLAS_STATUS_INVALID_VALUE;
    }
    if (!dC || lddc < m) return ICLBLAS_STATUS_INVALID_VALUE;
    if (dA && (!B || ldb < rb || ldda < ra)) return ICLBLAS_STATUS_INVALID_VALUE;
    if (dB && (!A || lda < ra || lddb < rb)) return ICLBLAS_STATUS_INVALID_VALUE;
    if (!iclblasContext::valid(handle)) return ICLBLAS_STATUS_INVALID_VALUE;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        const float a = *alpha;
        if (h->memory_mode == ICLBLAS_MEMORY_DEVICE) {
            icl::sgemm_backward_dev(h, transa, transb, m, n, k, a, A, lda, B, ldb, dC, lddc, dA, ldda, dB, lddb);
        } else {
            const float* hA = dB ? stage_in<float>(h, kSlotA, A, ra, ca, lda, true) : nullptr;
            const float* hB = dA ? stage_in<float>(h, kSlotB, B, rb, cb, ldb, true) : nullptr;
            const float* hdC = stage_in<float>(h, kSlotC, dC, m, n, lddc, true);
            float* hdA = dA ? stage_in<float>(h, kSlotW0, dA, ra, ca, ldda, false) : nullptr;
            float* hdB = dB ? stage_in<float>(h, kSlotW1, dB, rb, cb, lddb, false) : nullptr;
            icl::sgemm_backward_dev(h, transa, transb, m, n, k, a, hA, lda, hB, ldb, hdC, lddc, hdA, ldda, hdB, lddb);
            if (dA) stage_out<float>(h, dA, hdA, ra, ca, ldda);
            if (dB) stage_out<float>(h, dB, hdB, rb, cb, lddb);
        }
        finish(h);
    });
}

ICLBLAS_API iclblasStatus_t iclblasStrsmBackward(iclblasHandle_t handle, iclblasSideMode_t side_,
                                                 iclblasFillMode_t uplo_, iclblasOperation_t trans_,
                                                 iclblasDiagType_t diag_, int m, int n, const float* alpha, float* A,
                                                 int lda, float* X, int ldx, float* dX, int lddx, float* dA, int ldda,
                                                 float* dB, int lddb)
{
    const int trans = enum_arg(trans_);
    const int side = enum_arg(side_);
    const int uplo = enum_arg(uplo_);
    const int diag = enum_arg(diag_);
    bool work = false;
    const iclblasStatus_t st = trsm_checks(handle, side, uplo, trans, diag, m, n, alpha, A, lda, X, ldx, &work);
    if (!work) return st;
    const int order = side == ICLBLAS_SIDE_LEFT ? m : n;
    if (!dX || lddx < m) return ICLBLAS_STATUS_INVALID_VALUE;
    if (dA && ldda < order) return ICLBLAS_STATUS_INVALID_VALUE;
    if (dB && lddb < m) return ICLBLAS_STATUS_INVALID_VALUE;
    return guarded([&] {
        HandleLock lock(handle);
        iclblasContext* h = handle;
        const float a = *alpha;
        float* G = static_cast<float*>(h->scratch[kSlotW2].get((size_t)m * n * sizeof(float), h->stream));
        if (h->memory_mode == ICLBLAS_MEMORY_DEVICE) {
            icl::strsm_backward_dev(h, side, uplo, trans, diag, m, n, a, A, lda, X, ldx, dX, lddx, dA, ldda, dB, lddb,
                                    G);
        } else {
            const float* hA = stage_in<float>(h, kSlotA, A, order, order, lda, true);
            const float* hX = stage_in<float>(h, kSlotB, X, m, n, ldx, true);
            const float* hdX = stage_in<float>(h, kSlotC, dX, m, n, lddx, true);
            float* hdA = dA ? stage_in<float>(h, kSlotW0, dA, order, order, ldda, false) : nullptr;
            float* hdB = dB ? stage_in<float>(h, kSlotW1, dB, m, n, lddb, false) : nullptr;
            icl::strsm_backward_dev(h, side, uplo, trans, diag, m, n, a, hA, lda, hX, ldx, hdX, lddx, hdA, ldda, hdB,
                                    lddb, G);
            if (dA) stage_out<float>(h, dA, hdA, order, order, ldda);
            if (dB) stage_out<float>(h, dB, hdB, m, n, lddb);
        }
        finish(h);
        if (h->memory_mode == ICLBLAS_MEMORY_HOST) icl::trsm_check_error(h);
    });
}

} // extern "C"
