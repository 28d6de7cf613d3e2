// Free-list of device blocks (pool.hpp).
#include "pool.hpp"

#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace mlgpu {
namespace {

struct Block {
    void* p;
    size_t size;
    int device;
};

// free blocks beyond this many bytes per device are released
constexpr size_t kMaxCachedBytes = size_t{8} << 30;

std::mutex g_mutex;
std::unordered_map<void*, Block> g_owned;          // every live block
std::map<int, std::multimap<size_t, void*>> g_free;  // per device: free blocks by size
std::map<int, size_t> g_free_bytes;

size_t round_size(size_t b) {
    // 256 B granules up to 1 MiB, then 1 MiB granules: similar requests share blocks
    if (b <= (size_t{1} << 20)) return (b + 255) & ~size_t{255};
    return (b + (size_t{1} << 20) - 1) & ~((size_t{1} << 20) - 1);
}

}  // namespace

hipError_t pool_get(size_t bytes, void** out) {
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const size_t want = round_size(bytes == 0 ? 1 : bytes);
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        auto& fl = g_free[dev];
        auto it = fl.lower_bound(want);
        // reuse a free block that is not more than twice as large
        if (it != fl.end() && it->first <= 2 * want) {
            *out = it->second;
            g_free_bytes[dev] -= it->first;
            fl.erase(it);
            return hipSuccess;
        }
    }
    void* p = nullptr;
    e = hipMalloc(&p, want);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_mutex);
    g_owned[p] = Block{p, want, dev};
    *out = p;
    return hipSuccess;
}

void pool_put(void* p) {
    if (!p) return;
    std::vector<void*> release;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        auto o = g_owned.find(p);
        if (o == g_owned.end()) return;
        const Block b = o->second;
        auto& fl = g_free[b.device];
        fl.emplace(b.size, p);
        size_t& cached = g_free_bytes[b.device];
        cached += b.size;
        // over the cap: give back the largest free blocks
        while (cached > kMaxCachedBytes && !fl.empty()) {
            auto last = std::prev(fl.end());
            cached -= last->first;
            release.push_back(last->second);
            g_owned.erase(last->second);
            fl.erase(last);
        }
    }
    for (void* q : release) (void)hipFree(q);
}

}  // namespace mlgpu
