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

// Wave-size facts for CDNA4.
constexpr int WAVE = 64;
