// hip_emu.hpp -- CPU emulation of the HIP subset used by libcns (test
// infrastructure).  Workgroups run one after another; the lanes of a
// workgroup run as fibers on one host thread (or, with EMU_THREADS, as
// std::threads with a std::barrier), so __shared__ can be a function-local
// static.  Device memory is host memory, which lets AddressSanitizer check
// every kernel access on the CPU.
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
#include <ucontext.h>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__
#define __align__(n) alignas(n)
// a device symbol is an ordinary host object on the CPU
#define HIP_SYMBOL(sym) (&(sym))

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

inline void __syncthreads();

// raw buffer accesses: a resource is its base address; the offsets are the
// unsigned 32-bit byte offsets the hardware adds to it
struct __amdgpu_buffer_rsrc_t {
    char *base = nullptr;
};
struct emu_u32x2 {
    unsigned x, y;
};
inline __amdgpu_buffer_rsrc_t __builtin_amdgcn_make_buffer_rsrc(void *p, short, int, int) {
    __amdgpu_buffer_rsrc_t r;
    r.base = static_cast<char *>(p);
    return r;
}
inline emu_u32x2 __builtin_amdgcn_raw_buffer_load_b64(__amdgpu_buffer_rsrc_t r, int vo, int so, int) {
    emu_u32x2 v;
    std::memcpy(&v, r.base + (size_t)(unsigned)vo + (size_t)(unsigned)so, 8);
    return v;
}
inline unsigned __builtin_amdgcn_raw_buffer_load_b32(__amdgpu_buffer_rsrc_t r, int vo, int so, int) {
    unsigned v;
    std::memcpy(&v, r.base + (size_t)(unsigned)vo + (size_t)(unsigned)so, 4);
    return v;
}
inline void __builtin_amdgcn_raw_buffer_store_b64(emu_u32x2 v, __amdgpu_buffer_rsrc_t r, int vo, int so, int) {
    std::memcpy(r.base + (size_t)(unsigned)vo + (size_t)(unsigned)so, &v, 8);
}
inline void __builtin_amdgcn_raw_buffer_store_b32(unsigned v, __amdgpu_buffer_rsrc_t r, int vo, int so, int) {
    std::memcpy(r.base + (size_t)(unsigned)vo + (size_t)(unsigned)so, &v, 4);
}

// hardware reciprocal: IEEE division on the CPU
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
// hardware base-2 logarithm and exponential: the C library's on the CPU
inline float __builtin_amdgcn_logf(float x) { return std::log2(x); }
inline float __builtin_amdgcn_exp2f(float x) { return std::exp2(x); }

// scheduling hints and wave-uniform reads have no effect on the CPU
#define __builtin_amdgcn_sched_barrier(m) ((void)(m))
#define __builtin_amdgcn_s_setprio(p) ((void)(p))
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }

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
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
#define hipStreamNonBlocking 1

inline const char *hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) {
    *p = const_cast<void *>(sym);
    return hipSuccess;
}
inline hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
// a nominal 288 GB device with nothing allocated
inline hipError_t hipMemGetInfo(size_t *free_b, size_t *total_b) {
    *free_b = *total_b = (size_t)288 << 30;
    return hipSuccess;
}
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

#ifdef EMU_THREADS
// Persistent worker pool: one host thread per workgroup lane, reused for
// every workgroup of every launch (used by the AddressSanitizer build).
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
inline void __syncthreads() { emu_bar->arrive_and_wait(); }

template <class F, class... Args>
void emu_launch_now(F f, dim3 grid, dim3 block, Args... args) {
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
#else
// Fibers: every lane of a workgroup is a ucontext on one host thread;
// __syncthreads switches back to the scheduler, which resumes the lanes in
// order once all of them have arrived.  Lanes that return early leave the
// barrier count, as waves that end do on the hardware.
struct EmuFibers {
    std::vector<ucontext_t> ctx;
    std::vector<char *> stacks;
    std::vector<int> state;  // 0 runnable, 1 at barrier, 2 finished
    ucontext_t sched;
    unsigned cur = 0, nt = 0;
    std::function<void(unsigned)> task;
    static constexpr size_t STACK = 256 * 1024;
    void ensure(unsigned n) {
        while (stacks.size() < n) stacks.push_back((char *)std::malloc(STACK));
        if (ctx.size() < n) ctx.resize(n);
        state.assign(n, 0);
    }
    ~EmuFibers() {
        for (char *p : stacks) std::free(p);
    }
};
inline EmuFibers emu_fib;
inline dim3 emu_grid, emu_block;
inline unsigned emu_bx, emu_by, emu_bz;

inline void emu_set_ids(unsigned t) {
    threadIdx.x = t % emu_block.x;
    threadIdx.y = (t / emu_block.x) % emu_block.y;
    threadIdx.z = t / (emu_block.x * emu_block.y);
    blockIdx.x = emu_bx;
    blockIdx.y = emu_by;
    blockIdx.z = emu_bz;
    blockDim.x = emu_block.x;
    blockDim.y = emu_block.y;
    blockDim.z = emu_block.z;
    gridDim.x = emu_grid.x;
    gridDim.y = emu_grid.y;
    gridDim.z = emu_grid.z;
}
inline void emu_fiber_entry() {
    const unsigned t = emu_fib.cur;
    emu_fib.task(t);
    emu_fib.state[t] = 2;
    swapcontext(&emu_fib.ctx[t], &emu_fib.sched);
}
inline void __syncthreads() {
    const unsigned t = emu_fib.cur;
    emu_fib.state[t] = 1;
    swapcontext(&emu_fib.ctx[t], &emu_fib.sched);
}

template <class F, class... Args>
void emu_launch_now(F f, dim3 grid, dim3 block, Args... args) {
    const unsigned nt = block.x * block.y * block.z;
    emu_grid = grid;
    emu_block = block;
    emu_fib.ensure(nt);
    emu_fib.nt = nt;
    emu_fib.task = [&](unsigned t) { f(args...); };
    for (emu_bz = 0; emu_bz < grid.z; emu_bz++)
        for (emu_by = 0; emu_by < grid.y; emu_by++)
            for (emu_bx = 0; emu_bx < grid.x; emu_bx++) {
                for (unsigned t = 0; t < nt; t++) {
                    getcontext(&emu_fib.ctx[t]);
                    emu_fib.ctx[t].uc_stack.ss_sp = emu_fib.stacks[t];
                    emu_fib.ctx[t].uc_stack.ss_size = EmuFibers::STACK;
                    emu_fib.ctx[t].uc_link = nullptr;
                    makecontext(&emu_fib.ctx[t], (void (*)())emu_fiber_entry, 0);
                    emu_fib.state[t] = 0;
                }
                for (;;) {
                    bool any = false;
                    for (unsigned t = 0; t < nt; t++) {
                        if (emu_fib.state[t] == 2) continue;
                        any = true;
                        emu_fib.cur = t;
                        emu_fib.state[t] = 0;
                        emu_set_ids(t);
                        swapcontext(&emu_fib.sched, &emu_fib.ctx[t]);
                    }
                    if (!any) break;
                }
            }
}
#endif

// Stream capture: launches made while capturing are recorded and replayed
// by hipGraphLaunch (no other operation is captured by the product code).
struct EmuGraph {
    std::vector<std::function<void()>> ops;
};
typedef EmuGraph *hipGraph_t;
typedef EmuGraph *hipGraphExec_t;
enum { hipStreamCaptureModeThreadLocal = 2 };
inline thread_local EmuGraph *emu_capturing = nullptr;
inline hipError_t hipStreamBeginCapture(hipStream_t, int) {
    emu_capturing = new EmuGraph();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    *g = emu_capturing;
    emu_capturing = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t) {
    *e = g;
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto &op : e->ops) op();
    return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
struct hipDeviceProp_t {
    int multiProcessorCount;
};
// few "CUs" so that persistent launches give every workgroup several items
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
    p->multiProcessorCount = 8;
    return hipSuccess;
}

template <class F, class... Args>
void emu_launch(F f, dim3 grid, dim3 block, Args... args) {
    if (emu_capturing)
        emu_capturing->ops.push_back([=]() { emu_launch_now(f, grid, block, args...); });
    else
        emu_launch_now(f, grid, block, args...);
}

#define hipLaunchKernelGGL(F, G, B, SH, ST, ...) emu_launch(F, dim3(G), dim3(B), __VA_ARGS__)
