// Device memory for the working buffers of one entry-point call: blocks are
// borrowed from a per-process free list (allocated on first need, kept for
// reuse) and returned when the call ends.  Every entry point that borrows
// synchronises its stream before it returns the block, so a returned block
// is idle; concurrent calls (other threads, other streams) get distinct
// blocks.  hipMalloc / hipFree on every call would cost a device-wide
// synchronisation each.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace mlgpu {

hipError_t pool_get(size_t bytes, void** out);
void pool_put(void* p);

// a borrowed block returned at the end of its scope
class PoolBlock {
 public:
    PoolBlock() = default;
    PoolBlock(const PoolBlock&) = delete;
    PoolBlock& operator=(const PoolBlock&) = delete;
    ~PoolBlock() { release(); }
    hipError_t get(size_t bytes) {
        release();
        return pool_get(bytes, &p_);
    }
    void release() {
        if (p_) pool_put(p_);
        p_ = nullptr;
    }
    template <typename T = char>
    T* as() const { return static_cast<T*>(p_); }

 private:
    void* p_ = nullptr;
};

}  // namespace mlgpu
