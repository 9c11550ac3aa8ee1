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
__host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, T s) {
    return {a.re * s, a.im * s};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator*(cpx<T> a, T s) {
    return {a.re * s, a.im * s};
}
// a + b * s
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cfma(cpx<T> b, T s, cpx<T> a) {
    return {a.re + b.re * s, a.im + b.im * s};
}
// i * alpha * a
template <typename T>
__host__ __device__ __forceinline__ cpx<T> ialpha(T alpha, cpx<T> a) {
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

// Workgroups go to the 8 XCDs of the chip in turn (workgroup b to XCD
// b mod 8), and each XCD has its own L2.  This maps the workgroup index to
// a work index such that every XCD takes one contiguous range of the work
// (a bijection of [0, gridDim.x)): consecutive element-modes of one
// element and their neighbours, which share GLL points, meet in one L2.
__device__ __forceinline__ int xcd_block() {
    constexpr int NXCD = 8;
    const int b = (int)blockIdx.x;
    const int n = (int)gridDim.x;
    const int q = n / NXCD, rem = n - q * NXCD;
    const int x = b % NXCD, r = b / NXCD;
    return x * q + (x < rem ? x : rem) + r;
}

// Row stride (complex values) of an element's rows in the cardinal
// scratch: the nu1 orders of a row padded to a multiple of 16, so that
// every row starts on a 128-byte line and a run of 16 orders read or
// written by 16 lanes never straddles two lines.
__host__ __device__ __forceinline__ int scr_row(int nu1) {
    return (nu1 + 15) & ~15;
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
