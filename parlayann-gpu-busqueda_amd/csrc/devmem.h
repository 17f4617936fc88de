// devmem.h -- scoped device allocations and error propagation for the GPU graph builders.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace pann {

// Every allocation made through alloc() is freed when the object goes out of scope.
struct DevMem {
    std::vector<void*> ptrs;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T** p, size_t count) {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(q);
        *p = static_cast<T*>(q);
        return e;
    }
};

}  // namespace pann

#define BUILD_TRY(x)                        \
    do {                                    \
        const hipError_t e_ = (x);          \
        if (e_ != hipSuccess) return e_;    \
    } while (0)
