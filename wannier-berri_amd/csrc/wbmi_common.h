// Internal helpers shared by the libwbmi HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/wbmi.h"

struct wbmi_ctx {
    int device;
    hipStream_t stream;
    std::string err;
    void* eigh_scratch;  // this device's address of the no-allocation eigensolver's scratch
};

// Device address of the eigensolver scratch on the current device (eigh.hip); 0 on success.
int wbmi_eigh_scratch_init(void** ptr);

// Record an error message on the context and return the code.
int wbmi_fail(wbmi_ctx* ctx, int code, const char* fmt, ...);

#define WBMI_HIP_CHECK(ctx, call)                                                       \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return wbmi_fail((ctx), WBMI_ERR_HIP, "%s failed: %s (%s:%d)", #call,       \
                             hipGetErrorString(e_), __FILE__, __LINE__);                \
    } while (0)

#define WBMI_LAUNCH_CHECK(ctx) WBMI_HIP_CHECK(ctx, hipGetLastError())

// Orders one wave's own LDS accesses between phases (LDS written by some lanes
// and then read by other lanes of the same wave): the lanes of a wave execute
// together, so no s_barrier is needed; the fences wait for the wave's
// outstanding LDS operations and keep the compiler from moving accesses across.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Complex number with natural alignment for vector loads.
template <typename T>
struct __attribute__((aligned(2 * sizeof(T)))) cplx {
    T re, im;
};

template <typename T>
__host__ __device__ inline cplx<T> mk(T r, T i) { cplx<T> z; z.re = r; z.im = i; return z; }
template <typename T>
__host__ __device__ inline cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T>
__host__ __device__ inline cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T>
__host__ __device__ inline cplx<T> operator-(cplx<T> a) { return mk(-a.re, -a.im); }
template <typename T>
__host__ __device__ inline cplx<T> operator*(cplx<T> a, cplx<T> b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ inline cplx<T> operator*(T s, cplx<T> a) { return mk(s * a.re, s * a.im); }
template <typename T>
__host__ __device__ inline cplx<T> conj(cplx<T> a) { return mk(a.re, -a.im); }
// a * conj(b)
template <typename T>
__host__ __device__ inline cplx<T> mulc(cplx<T> a, cplx<T> b) {
    return mk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
// conj(a) * b
template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return mk(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}
template <typename T>
__host__ __device__ inline cplx<T> times_i(cplx<T> a) { return mk(-a.im, a.re); }
template <typename T>
__host__ __device__ inline cplx<T>& operator+=(cplx<T>& a, cplx<T> b) { a.re += b.re; a.im += b.im; return a; }
// acc += a b and acc += conj(a) b as four fused multiply-adds (a product
// followed by an add would round twice and cost six operations)
template <typename T>
__host__ __device__ inline void cfma(cplx<T>& acc, cplx<T> a, cplx<T> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}
template <typename T>
__host__ __device__ inline void cfma_cj(cplx<T>& acc, cplx<T> a, cplx<T> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(-a.im, b.re, acc.im);
}
template <typename T>
__host__ __device__ inline cplx<T>& operator-=(cplx<T>& a, cplx<T> b) { a.re -= b.re; a.im -= b.im; return a; }

template <typename T>
__device__ inline cplx<double> to_c64(cplx<T> a) { return mk((double)a.re, (double)a.im); }

// Fill kernel (used instead of hipMemsetAsync so callers may capture graphs).
int wbmi_zero(wbmi_ctx* ctx, void* ptr, size_t bytes);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
