// CPU stand-in for <hip/hip_runtime.h>, used ONLY by the CPU emulation build
// of the engine (tests/cpu_emul/build.sh).  Device memory is host heap memory
// sized exactly as requested, so AddressSanitizer reports any out-of-bounds
// access a kernel makes; kernels run block by block, thread by thread.
//
// Thread-to-thread communication (__shfl_xor, __syncthreads, shared memory)
// is NOT emulated: values computed through it are meaningless here.  The
// build checks memory safety of every global access, not numerical results.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define __global__
#define __device__
#define __host__
#define __shared__
#define __launch_bounds__(...)
#define __forceinline__ inline
#define __align__(x) __attribute__((aligned(x)))
#define HIP_KERNEL_NAME(...) __VA_ARGS__

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct sq_emul_idx { unsigned x = 0, y = 0, z = 0; };
inline thread_local sq_emul_idx threadIdx, blockIdx, blockDim, gridDim;

inline void __syncthreads() {}
inline int __shfl_xor(int v, int, int = 64) { return v; }
inline double __shfl_xor(double v, int, int = 64) { return v; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline double __dadd_rn(double a, double b) { return a + b; }
inline int __builtin_amdgcn_readfirstlane(int x) { return x; }
// Cross-lane permutations are not modelled (one lane runs at a time): the
// value stays with its lane, like __shfl_xor above.
inline int __builtin_amdgcn_update_dpp(int, int x, int, int, int, bool) { return x; }
inline int __builtin_amdgcn_mov_dpp(int x, int, int, int, bool) { return x; }
typedef unsigned sq_emul_u2 __attribute__((ext_vector_type(2)));
inline sq_emul_u2 __builtin_amdgcn_permlane16_swap(unsigned a, unsigned b, bool, bool) { return sq_emul_u2{a, b}; }
inline sq_emul_u2 __builtin_amdgcn_permlane32_swap(unsigned a, unsigned b, bool, bool) { return sq_emul_u2{a, b}; }
inline double __dsub_rn(double a, double b) { return a - b; }
inline double __dmul_rn(double a, double b) { return a * b; }

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum { hipStreamNonBlocking = 1 };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1, hipDeviceAttributeMaxSharedMemoryPerBlock = 2 };
struct sq_emul_stream {};
struct sq_emul_event {};
typedef sq_emul_stream *hipStream_t;
typedef sq_emul_event *hipEvent_t;

inline const char *hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipMemGetInfo(size_t *free_b, size_t *total_b) {
    *free_b = (size_t)64 << 30;   // plenty: the emulation's states are small
    *total_b = (size_t)64 << 30;
    return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int) {
    *v = a == hipDeviceAttributeMultiprocessorCount ? 4 : 65536;
    return hipSuccess;
}
inline hipError_t hipMalloc(void **p, size_t n) {
    *p = std::malloc(n);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc(reinterpret_cast<void **>(p), n); }
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
enum { hipHostMallocDefault = 0 };
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned) { return hipMalloc(p, n); }
inline hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t) {
    return hipMemcpy(d, s, n, k);
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = new sq_emul_stream; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new sq_emul_event; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
// a __device__ array is an ordinary host array here: its address is itself
#define HIP_SYMBOL(x) (&(x))
template <typename S> inline hipError_t hipGetSymbolAddress(void **p, const S *sym) {
    *p = const_cast<void *>(static_cast<const void *>(sym));
    return hipSuccess;
}
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0; return hipSuccess; }

// Matrix-core instructions are wave-wide; here they leave the accumulator as is.
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) ((void)(a), (void)(b), (c))
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) ((void)(a), (void)(b), (c))
#define __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, x, y, z) ((void)(a), (void)(b), (c))

template <typename F, typename... A>
void sq_emul_launch(F kernel, dim3 grid, dim3 block, size_t, hipStream_t, A... args) {
    gridDim.x = grid.x;
    blockDim.x = block.x;
    for (unsigned b = 0; b < grid.x; b++) {
        blockIdx.x = b;
        for (unsigned t = 0; t < block.x; t++) {
            threadIdx.x = t;
            kernel(args...);
        }
    }
}
#define hipLaunchKernelGGL(K, G, B, L, S, ...) sq_emul_launch(K, dim3(G), dim3(B), (size_t)(L), S, __VA_ARGS__)
