// Shared device/host definitions for the bfps HIP backend (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define BFPSH_DEV __device__ __forceinline__

template <typename T> struct cplx_t;
template <> struct cplx_t<float> { typedef float2 type; };
template <> struct cplx_t<double> { typedef double2 type; };
template <typename T> using C2 = typename cplx_t<T>::type;

template <typename T> BFPSH_DEV C2<T> mk(T re, T im) { C2<T> r; r.x = re; r.y = im; return r; }
template <typename T> BFPSH_DEV C2<T> cadd(C2<T> a, C2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> BFPSH_DEV C2<T> csub(C2<T> a, C2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> BFPSH_DEV C2<T> cmul(C2<T> a, C2<T> b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> BFPSH_DEV C2<T> cconj(C2<T> a) { return mk<T>(a.x, -a.y); }
// multiply by (sign * i): sign = -1 forward (multiply by -i), +1 backward (by +i)
template <typename T, int SIGN> BFPSH_DEV C2<T> cmul_i(C2<T> a) {
    return SIGN < 0 ? mk<T>(a.y, -a.x) : mk<T>(-a.y, a.x);
}

// Device layout strides.  A spectral plane (one component) holds N ky rows,
// each the kz-kx slab of N*N/2 complex values; consecutive ky rows are
// row_stride(N) = N*N/2 + BFPSH_PAD_Y complex apart and the components of
// a vector field plane_size(N) = N*row_stride(N) apart.  The pad breaks
// the power-of-two stride of the y passes (4 MiB at 1024^3 FP32), whose
// rows otherwise meet on the same HBM banks: one strided y pass over a
// plane measured 2.1 -> 2.9 TB/s at 1024^3 FP32 and 4.5 -> 4.8 TB/s at
// 512^3 with a pad of 16.  A real plane uses the same bytes: real row y
// (the [z][x] slab) starts at y * 2*row_stride(N) reals.  Pad elements are
// never read as data.
#ifndef BFPSH_PAD_Y
#define BFPSH_PAD_Y 16
#endif
__host__ __device__ constexpr size_t row_stride(int n) { return size_t(n) * size_t(n / 2) + BFPSH_PAD_Y; }
__host__ __device__ constexpr size_t plane_size(int n) { return size_t(n) * row_stride(n); }
// physical offset (complex) of the logical spectral index (ky*N + kz)*N/2 + kx
// (logical indices of N <= 1024 fit 32 bits)
__host__ __device__ inline size_t sphys(int n, size_t idx) {
    const uint32_t s = uint32_t(n) * uint32_t(n / 2), i = uint32_t(idx);
    return size_t(i / s) * row_stride(n) + i % s;
}

// Wave-size facts for CDNA4.
constexpr int WAVE = 64;

// Cross-lane moves inside rows of 16 lanes (DPP, VALU only: no LDS traffic).
// CTRL is a DPP control code: 0x150 + k broadcasts lane k of each row to the
// whole row (row_newbcast), 0x101..0x10F shift the row by 1..15 towards
// lane 0 (row_shl: lane i reads lane i + s; lanes past the row end read 0),
// 0x121..0x12F rotate it (row_ror).
template <int CTRL> BFPSH_DEV float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL> BFPSH_DEV double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
// sum over the 16 lanes of each row, left in every lane of the row
BFPSH_DEV double row16_sum(double v) {
    v += dpp_d<0x128>(v);
    v += dpp_d<0x124>(v);
    v += dpp_d<0x122>(v);
    v += dpp_d<0x121>(v);
    return v;
}
