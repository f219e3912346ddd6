// Pinned staging of the host-pointer memory mode (see PinnedStager in
// handle.hpp). The reference creates a device buffer per argument per call and
// maps or copies host memory (core/src/ocl/ocl_buffer.cpp:27-111); here the
// device side is the handle's grow-only scratch and the host side goes through
// two reused pinned buffers, so a call costs one pass over PCIe per matrix at
// the pinned rate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "handle.hpp"

namespace icl {

namespace {

constexpr size_t kChunk = 32u << 20; // bytes per pipeline stage
constexpr size_t kParallelMin = 4u << 20;

int copy_threads()
{
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(8u, hc ? hc : 1u));
}

// memcpy of one contiguous range, split over a few threads when large.
void copy_bytes(void* dst, const void* src, size_t bytes)
{
    const int nt = bytes >= kParallelMin ? copy_threads() : 1;
    if (nt == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    std::vector<std::thread> ts;
    const size_t per = (bytes + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const size_t o = (size_t)t * per;
        if (o >= bytes) break;
        const size_t len = std::min(per, bytes - o);
        ts.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + o, static_cast<const char*>(src) + o, len); });
    }
    for (auto& t : ts) t.join();
}

// Columns [0, cols) of a rows x cols column-major block from src (pitch ld
// elements) into dst (pitch ld), rows only, split over threads by columns.
void copy_columns(char* dst, const char* src, int rows, int cols, long ld, size_t es)
{
    const size_t col_bytes = (size_t)rows * es, pitch = (size_t)ld * es;
    const int nt = col_bytes * cols >= kParallelMin ? std::min(copy_threads(), cols) : 1;
    auto work = [=](int c0, int c1) {
        for (int c = c0; c < c1; ++c) std::memcpy(dst + c * pitch, src + c * pitch, col_bytes);
    };
    if (nt == 1) {
        work(0, cols);
        return;
    }
    std::vector<std::thread> ts;
    const int per = (cols + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int c0 = t * per, c1 = std::min(cols, c0 + per);
        if (c0 >= c1) break;
        ts.emplace_back(work, c0, c1);
    }
    for (auto& t : ts) t.join();
}

} // namespace

void PinnedStager::ensure(size_t bytes)
{
    if (bytes <= cap_ && buf_[0]) return;
    release();
    for (int i = 0; i < 2; ++i) {
        hip_check(hipHostMalloc(&buf_[i], bytes, hipHostMallocDefault), "hipHostMalloc");
        hip_check(hipEventCreateWithFlags(&done_[i], hipEventDisableTiming), "hipEventCreate");
    }
    cap_ = bytes;
}

void PinnedStager::release()
{
    for (int i = 0; i < 2; ++i) {
        if (done_[i]) (void)hipEventSynchronize(done_[i]);
        if (buf_[i]) (void)hipHostFree(buf_[i]);
        if (done_[i]) (void)hipEventDestroy(done_[i]);
        buf_[i] = nullptr;
        done_[i] = nullptr;
    }
    cap_ = 0;
}

void PinnedStager::h2d(void* dst, const void* src, size_t bytes, hipStream_t stream)
{
    if (bytes == 0) return;
    ensure(kChunk);
    size_t off = 0;
    for (int i = 0; off < bytes; ++i) {
        const int b = i & 1;
        const size_t len = std::min(kChunk, bytes - off);
        // the DMA that last read buf_[b] must be done before it is refilled
        hip_check(hipEventSynchronize(done_[b]), "hipEventSynchronize");
        copy_bytes(buf_[b], static_cast<const char*>(src) + off, len);
        hip_check(hipMemcpyAsync(static_cast<char*>(dst) + off, buf_[b], len, hipMemcpyHostToDevice, stream),
                  "hipMemcpyAsync(H2D)");
        hip_check(hipEventRecord(done_[b], stream), "hipEventRecord");
        off += len;
    }
}

void PinnedStager::d2h(void* dst, const void* src, int rows, int cols, long ld, size_t es, hipStream_t stream)
{
    if (rows <= 0 || cols <= 0) return;
    ensure(kChunk);
    const size_t pitch = (size_t)ld * es;
    // Whole columns per chunk (at least one; a column longer than a chunk
    // grows the buffers).
    const size_t col_bytes = (size_t)rows * es;
    int per = col_bytes >= kChunk ? 1 : (int)std::min<size_t>((size_t)cols, (kChunk - col_bytes) / pitch + 1);
    per = std::max(1, std::min(per, cols));
    const size_t need = (size_t)(per - 1) * pitch + (size_t)rows * es;
    if (need > cap_) ensure(need);
    const int chunks = (cols + per - 1) / per;
    auto enqueue = [&](int i) {
        const int b = i & 1, c0 = i * per, nc = std::min(per, cols - c0);
        const size_t len = (size_t)(nc - 1) * pitch + (size_t)rows * es;
        hip_check(hipEventSynchronize(done_[b]), "hipEventSynchronize"); // buffer free (earlier use drained)
        hip_check(hipMemcpyAsync(buf_[b], static_cast<const char*>(src) + c0 * pitch, len, hipMemcpyDeviceToHost,
                                 stream),
                  "hipMemcpyAsync(D2H)");
        hip_check(hipEventRecord(done_[b], stream), "hipEventRecord");
    };
    enqueue(0);
    for (int i = 0; i < chunks; ++i) {
        if (i + 1 < chunks) enqueue(i + 1);
        const int b = i & 1, c0 = i * per, nc = std::min(per, cols - c0);
        hip_check(hipEventSynchronize(done_[b]), "hipEventSynchronize");
        copy_columns(static_cast<char*>(dst) + c0 * pitch, static_cast<const char*>(buf_[b]), rows, nc, ld, es);
    }
}

} // namespace icl
