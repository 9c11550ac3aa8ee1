// axs3d_common.hpp -- shared device types and helpers (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axs3d {

constexpr int NPED = 5;    // GLL points per direction (npol = 4)
constexpr int NPEM = 25;   // GLL points per element

// Complex value; aligned to its size so that one LDS or global access
// moves both parts.
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, T s) {
    return {a.re * s, a.im * s};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename T>
__device__ __forceinline__ cpx<T> operator*(cpx<T> a, T s) {
    return {a.re * s, a.im * s};
}
// a + b * s
template <typename T>
__device__ __forceinline__ cpx<T> cfma(cpx<T> b, T s, cpx<T> a) {
    return {a.re + b.re * s, a.im + b.im * s};
}
// i * alpha * a
template <typename T>
__device__ __forceinline__ cpx<T> ialpha(T alpha, cpx<T> a) {
    return {-alpha * a.im, alpha * a.re};
}

// Orders the LDS accesses of one wave: every lane's accesses before the
// call complete before any lane's accesses after it (LDS executes a
// wave's instructions in order; the fences keep the compiler from moving
// accesses across).  For data exchanged only within one wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Element-mode work item: element index and Fourier order alpha.
struct EMItem {
    int32_t elem;
    int32_t alpha;
};

// Element operand rows in the device layout [e][ELEM_OPS][25].
enum ElemOpRow {
    OP_DSDXII = 0, OP_DSDETA, OP_DZDXII, OP_DZDETA, OP_INVS,
    OP_DSDXII_IF, OP_DSDETA_IF, OP_DZDXII_IF, OP_DZDETA_IF, OP_INVS_IF,
    ELEM_OPS
};

}  // namespace axs3d
