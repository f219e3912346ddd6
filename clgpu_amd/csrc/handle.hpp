// Library context behind iclblasHandle_t.
//
// Replaces the reference's iclblasContext + iclgpu::context + OpenCL toolkit
// (libs/blas/shlib/iclBLASImpl.hpp:25-35, libs/blas/shlib/iclBLAS.cpp:21-39,
// core/src/ocl/ocl_toolkit.cpp:92-191): instead of an OpenCL context, a queue and
// a JIT program cache, a handle holds one HIP device, one HIP stream, the
// memory/math modes, an implementation override and grow-only device scratch
// buffers used to stage host matrices and hold refinement temporaries.
#pragma once
#include <hip/hip_runtime_api.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "iclBLAS.h"
#include "iclBLAS_ext.h"
#include "kernel_abi.h"

namespace icl {

// Error carrying an iclblasStatus_t across the C++ layer; converted back to a
// status at the C boundary (nothing is thrown across extern "C").
struct status_error : std::runtime_error {
    iclblasStatus_t status;
    status_error(iclblasStatus_t s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw status_error(ICLBLAS_STATUS_ALLOC_FAILED, std::string(what) + ": " + hipGetErrorString(e));
    throw status_error(ICLBLAS_STATUS_EXECUTION_FAILED, std::string(what) + ": " + hipGetErrorString(e));
}

// Device memory held by a handle's scratch buffers, and its cap
// (iclblasSetScratchLimit; 0 = none).
struct ScratchBudget {
    size_t limit = 0;
    size_t held = 0;
};

// A grow-only device allocation owned by the handle.
class DeviceScratch {
public:
    DeviceScratch() = default;
    DeviceScratch(const DeviceScratch&) = delete;
    DeviceScratch& operator=(const DeviceScratch&) = delete;
    ~DeviceScratch() { release(); }

    void set_budget(ScratchBudget* b) { budget_ = b; }

    // Returns at least `bytes` of device memory. Growing waits for `stream` first
    // because queued work may still use the old allocation. Growing beyond the
    // handle's cap fails like an exhausted device (ALLOC_FAILED), before anything
    // is freed or allocated.
    void* get(size_t bytes, hipStream_t stream)
    {
        if (bytes <= cap_ && ptr_) return ptr_;
        const size_t want = bytes < 256 ? 256 : bytes;
        if (budget_ && budget_->limit && budget_->held - cap_ + want > budget_->limit)
            throw status_error(ICLBLAS_STATUS_ALLOC_FAILED, "scratch limit of the handle");
        if (ptr_) {
            hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            release();
        }
        hip_check(hipMalloc(&ptr_, want), "hipMalloc");
        cap_ = want;
        if (budget_) budget_->held += want;
        return ptr_;
    }

    void release()
    {
        if (ptr_) (void)hipFree(ptr_);
        if (budget_) budget_->held -= cap_;
        ptr_ = nullptr;
        cap_ = 0;
    }

private:
    void* ptr_ = nullptr;
    size_t cap_ = 0;
    ScratchBudget* budget_ = nullptr;
};

// Host <-> device copies of the host-pointer memory mode through pinned
// memory. Pageable copies run at about half the PCIe rate, and pitched ones on
// pageable memory are split into one transfer per column, so instead the
// handle keeps two pinned buffers (hipHostMalloc, grown once and reused) and
// pipelines chunks: while the DMA engine moves one buffer, CPU threads fill or
// drain the other. Declared here, defined in staging.cpp.
class PinnedStager {
public:
    PinnedStager() = default;
    PinnedStager(const PinnedStager&) = delete;
    PinnedStager& operator=(const PinnedStager&) = delete;
    ~PinnedStager() { release(); }

    // dst (device) <- the `bytes` at src (host), enqueued on `stream`. Returns once
    // src may be reused (its bytes are in pinned memory or on the device).
    void h2d(void* dst, const void* src, size_t bytes, hipStream_t stream);
    // Column-major host matrix (rows x cols, leading dimension ld, es bytes per
    // element) <- the same layout on the device after the work queued on
    // `stream`; rows ld > rows of the host matrix are not written. Synchronous.
    void d2h(void* dst, const void* src, int rows, int cols, long ld, size_t es, hipStream_t stream);
    void release();

private:
    void ensure(size_t bytes);
    void* buf_[2] = {nullptr, nullptr};
    hipEvent_t done_[2] = {nullptr, nullptr}; // DMA of the chunk last placed in buf_[i] finished
    size_t cap_ = 0;
};

// A stream for host <-> device copies that overlap the kernels on the handle's
// stream (non-blocking, default priority), with events that order the two per
// pipeline step (reused round robin). Created on first use.
struct CopyLane {
    static constexpr int kEvents = 16;
    hipStream_t stream = nullptr;
    hipEvent_t in[kEvents] = {}, out[kEvents] = {};
    CopyLane() = default;
    CopyLane(const CopyLane&) = delete;
    CopyLane& operator=(const CopyLane&) = delete;
    ~CopyLane() { release(); }

    void ensure()
    {
        if (stream) return;
        for (int i = 0; i < kEvents; ++i) {
            hip_check(hipEventCreateWithFlags(&in[i], hipEventDisableTiming), "hipEventCreate");
            hip_check(hipEventCreateWithFlags(&out[i], hipEventDisableTiming), "hipEventCreate");
        }
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    }

    void release()
    {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
            stream = nullptr;
        }
        for (int i = 0; i < kEvents; ++i) {
            if (in[i]) (void)hipEventDestroy(in[i]);
            if (out[i]) (void)hipEventDestroy(out[i]);
            in[i] = out[i] = nullptr;
        }
    }
};

// The side stream of a handle's large split-BF16 GEMMs (icl_side_lane): the
// second half of the operand prepass runs there while the first products run
// on the handle's stream. Non-blocking, default priority; created on first use.
struct SideLane {
    icl_side_lane lane{};
    SideLane() = default;
    SideLane(const SideLane&) = delete;
    SideLane& operator=(const SideLane&) = delete;
    ~SideLane() { release(); }

    const icl_side_lane* ensure()
    {
        if (lane.stream) return &lane;
        hip_check(hipEventCreateWithFlags(&lane.start, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&lane.done, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipStreamCreateWithFlags(&lane.stream, hipStreamNonBlocking), "hipStreamCreate");
        return &lane;
    }

    void release()
    {
        if (lane.stream) {
            (void)hipStreamSynchronize(lane.stream);
            (void)hipStreamDestroy(lane.stream);
        }
        if (lane.start) (void)hipEventDestroy(lane.start);
        if (lane.done) (void)hipEventDestroy(lane.done);
        lane = icl_side_lane{};
    }
};

// Opt-in timing of the main kernel of each GEMM launch (iclblasSetKernelTiming):
// a begin/end HIP event pair recorded on the launch stream around that kernel
// alone. Events are created on first use and reused; at most kMaxEvents are
// recorded between two reads (later marks are counted as dropped). They are
// created without the system-scope fence: they only measure time (the read
// synchronizes the stream first), and a fenced record writes back and
// invalidates the caches between the kernels it brackets, which adds about
// 50 us to a split-BF16 8192^3 call.
struct KernelTimer {
    static constexpr size_t kMaxEvents = 8192;
    bool on = false;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    long dropped = 0;
    KernelTimer() = default;
    KernelTimer(const KernelTimer&) = delete;
    KernelTimer& operator=(const KernelTimer&) = delete;
    ~KernelTimer() { release(); }

    hipError_t mark(hipStream_t s)
    {
        if (used == kMaxEvents) {
            ++dropped;
            return hipSuccess;
        }
        if (used == ev.size()) {
            hipEvent_t e = nullptr;
            const hipError_t err = hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
            if (err != hipSuccess) return err;
            ev.push_back(e);
        }
        return hipEventRecord(ev[used++], s);
    }

    void release()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        used = 0;
        dropped = 0;
    }
};

// The timer of the handle whose call the current thread is running (set for the
// duration of a call when that handle's timing is on), read by icl_timer_mark.
extern thread_local KernelTimer* t_active_timer;
// The side lane offered to the GEMM launch the current thread is running (set by
// run_gemm for the duration of one launch), read by icl_gemm_side_lane.
extern thread_local const icl_side_lane* t_side_lane;

} // namespace icl

struct iclblasContext {
    static constexpr int kTag = 0xB1A5;
    static constexpr int kScratchSlots = 12;
    static constexpr int kSlotTrsmWorkspace = 6; // inverted diagonal blocks + solution of blocked STRSM
    static constexpr int kSlotGemmWorkspace = 7; // packed operands of split-BF16 GEMMs
    static constexpr int kSlotExpanded = 8;      // dense copy of a triangular / symmetric operand (level 3)
    static constexpr int kSlotLevel3Out = 9;     // TRMM result when C overlaps B in device memory mode
    static constexpr int kSlotGemvPartial = 10;  // split partial sums of GEMV
    static constexpr int kSlotRefine = 11;       // per-block maxima of the refinement convergence test

    int tag = kTag;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    iclblasMemoryMode_t memory_mode = ICLBLAS_MEMORY_HOST;
    iclblasMathMode_t math_mode = ICLBLAS_MATH_FP32;
    std::string forced_impl;
    std::string last_impl;
    int last_sweeps = 0; // refinement sweeps the last iclblasStrsmRefined ran
    std::mutex mu; // a handle serializes its own calls
    icl::ScratchBudget budget; // cap and total of the scratch buffers below
    icl::DeviceScratch scratch[kScratchSlots];
    icl::PinnedStager stager;
    icl::CopyLane copy;
    icl::SideLane side;
    int* trsm_error = nullptr; // device word set by an aborted dataflow STRSM (allocated on first use)
    bool trsm_error_pending = false; // a dataflow STRSM was queued since the word was last read
    icl::KernelTimer timer;
    // Caller-provided GEMM workspace (iclblasSgemmNoAlloc): while set, GEMMs
    // place their workspace here instead of in the grow-only scratch, and a
    // GEMM needing more than ext_ws_bytes fails with INVALID_VALUE.
    void* ext_ws = nullptr;
    size_t ext_ws_bytes = 0;

    static bool valid(const iclblasContext* h) { return h != nullptr && h->tag == kTag; }
};
