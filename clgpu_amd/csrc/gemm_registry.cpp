// Storage behind icl_register_gemm_variants / icl_gemm_variant_at: every GEMM
// kernel table registers itself while the library loads; readers see the
// entries sorted by name so that selection does not depend on link order.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "kernel_abi.h"

namespace {

struct Registry {
    std::mutex mu;
    std::vector<const icl_gemm_variant*> entries;
    bool sorted = false;
};

Registry& registry()
{
    static Registry r;
    return r;
}

void sort_once(Registry& r)
{
    if (r.sorted) return;
    std::sort(r.entries.begin(), r.entries.end(),
              [](const icl_gemm_variant* a, const icl_gemm_variant* b) { return std::strcmp(a->name, b->name) < 0; });
    r.sorted = true;
}

} // namespace

extern "C" void icl_register_gemm_variants(const icl_gemm_variant* table, int count)
{
    Registry& r = registry();
    std::lock_guard<std::mutex> g(r.mu);
    for (int i = 0; i < count; ++i) r.entries.push_back(table + i);
    r.sorted = false;
}

extern "C" int icl_gemm_variant_count(void)
{
    Registry& r = registry();
    std::lock_guard<std::mutex> g(r.mu);
    return (int)r.entries.size();
}

extern "C" const icl_gemm_variant* icl_gemm_variant_at(int index)
{
    Registry& r = registry();
    std::lock_guard<std::mutex> g(r.mu);
    sort_once(r);
    if (index < 0 || index >= (int)r.entries.size()) return nullptr;
    return r.entries[(size_t)index];
}
