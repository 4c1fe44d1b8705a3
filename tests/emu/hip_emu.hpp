// hip_emu.hpp -- CPU emulation of the HIP subset used by libcns (test
// infrastructure).  Workgroups run one after another; the threads of a
// workgroup run as std::threads with a std::barrier for __syncthreads, so
// __shared__ can be a function-local static.  Device memory is host memory,
// which lets AddressSanitizer check every kernel access on the CPU.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx {
    unsigned x = 0, y = 0, z = 0;
};
inline thread_local emu_idx threadIdx, blockIdx, blockDim, gridDim;
inline thread_local std::barrier<> *emu_bar = nullptr;
inline double emu_shfl[1024];

inline void __syncthreads() { emu_bar->arrive_and_wait(); }

inline double __shfl_xor(double v, int off) {
    const unsigned t = threadIdx.x;
    emu_shfl[t] = v;
    __syncthreads();
    double r = emu_shfl[t ^ (unsigned)off];
    __syncthreads();
    return r;
}
inline long long __double_as_longlong(double d) {
    long long l;
    std::memcpy(&l, &d, 8);
    return l;
}
inline double __longlong_as_double(long long l) {
    double d;
    std::memcpy(&d, &l, 8);
    return d;
}
inline unsigned long long atomicMax(unsigned long long *a, unsigned long long v) {
    unsigned long long old = __atomic_load_n(a, __ATOMIC_SEQ_CST);
    while (old < v && !__atomic_compare_exchange_n(a, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
using std::max;
using std::min;

typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
#define hipStreamNonBlocking 1

inline const char *hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) {
    *n = 1;
    return hipSuccess;
}
inline hipError_t hipMalloc(void **p, size_t b) {
    // poison new allocations with a NaN pattern so reads of unwritten memory show up
    *p = std::malloc(b ? b : 1);
    if (!*p) return hipErrorOutOfMemory;
    std::memset(*p, 0xff, b);
    return hipSuccess;
}
inline hipError_t hipFree(void *p) {
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t b, hipMemcpyKind) {
    std::memcpy(d, s, b);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t b, hipMemcpyKind k, hipStream_t) {
    return hipMemcpy(d, s, b, k);
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
    *s = (void *)1;
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) {
    *e = (void *)1;
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *f, hipEvent_t, hipEvent_t) {
    *f = 0.f;
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }

// Persistent worker pool: one host thread per workgroup lane, reused for
// every workgroup of every launch.
struct EmuPool {
    std::vector<std::thread> th;
    unsigned n = 0;
    std::function<void(unsigned)> task;
    unsigned active = 0;
    bool quit = false;
    std::barrier<> *start = nullptr, *done = nullptr;
    void ensure(unsigned want) {
        if (want <= n) return;
        shutdown();
        n = want;
        start = new std::barrier<>((std::ptrdiff_t)n + 1);
        done = new std::barrier<>((std::ptrdiff_t)n + 1);
        quit = false;
        for (unsigned t = 0; t < n; t++)
            th.emplace_back([this, t]() {
                for (;;) {
                    start->arrive_and_wait();
                    if (quit) break;
                    if (t < active) task(t);
                    done->arrive_and_wait();
                }
            });
    }
    void run(unsigned nt, std::function<void(unsigned)> f) {
        task = std::move(f);
        active = nt;
        start->arrive_and_wait();
        done->arrive_and_wait();
    }
    void shutdown() {
        if (!n) return;
        quit = true;
        start->arrive_and_wait();
        for (auto &x : th) x.join();
        th.clear();
        delete start;
        delete done;
        n = 0;
    }
    ~EmuPool() { shutdown(); }
};
inline EmuPool emu_pool;

template <class F, class... Args>
void emu_launch(F f, dim3 grid, dim3 block, Args... args) {
    const unsigned nt = block.x * block.y * block.z;
    emu_pool.ensure(nt < 64 ? 64 : nt);
    for (unsigned bz = 0; bz < grid.z; bz++)
        for (unsigned by = 0; by < grid.y; by++)
            for (unsigned bx = 0; bx < grid.x; bx++) {
                std::barrier<> bar((std::ptrdiff_t)nt);
                emu_pool.run(nt, [&](unsigned t) {
                    threadIdx.x = t % block.x;
                    threadIdx.y = (t / block.x) % block.y;
                    threadIdx.z = t / (block.x * block.y);
                    blockIdx.x = bx;
                    blockIdx.y = by;
                    blockIdx.z = bz;
                    blockDim.x = block.x;
                    blockDim.y = block.y;
                    blockDim.z = block.z;
                    gridDim.x = grid.x;
                    gridDim.y = grid.y;
                    gridDim.z = grid.z;
                    emu_bar = &bar;
                    f(args...);
                });
            }
}
#define hipLaunchKernelGGL(F, G, B, SH, ST, ...) emu_launch(F, dim3(G), dim3(B), __VA_ARGS__)
