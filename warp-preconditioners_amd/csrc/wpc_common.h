// Shared device helpers for the gfx950 kernels: storage formats, the binary16
// arithmetic model, wave64 primitives and launch plumbing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "wpc.h"

namespace wpc {

constexpr int kWave = 64;

// ------------------------------------------------------------------ storage
struct f16s { uint16_t b; };
struct bf16s { uint16_t b; };

__device__ __forceinline__ float f16_to_float(uint16_t b) {
    return (float)__builtin_bit_cast(_Float16, b);
}
__device__ __forceinline__ uint16_t float_to_f16(float f) {  // round to nearest even
    return __builtin_bit_cast(uint16_t, (_Float16)f);
}
__device__ __forceinline__ float bf16_to_float(uint16_t b) {
    return __uint_as_float(((uint32_t)b) << 16);
}
__device__ __forceinline__ uint16_t float_to_bf16(float f) {
    uint32_t bits = __float_as_uint(f);
    if (isnan(f)) return (uint16_t)((bits >> 16) | 0x40);
    bits += 0x7fffu + ((bits >> 16) & 1u);
    return (uint16_t)(bits >> 16);
}

// Load a stored element as a compute value of type C, and store back.
template <class S> struct St;
template <> struct St<double> {
    template <class C> __device__ static C ld(const double* p, long i) { return (C)p[i]; }
    template <class C> __device__ static void st(double* p, long i, C v) { p[i] = (double)v; }
};
template <> struct St<float> {
    template <class C> __device__ static C ld(const float* p, long i) { return (C)p[i]; }
    template <class C> __device__ static void st(float* p, long i, C v) { p[i] = (float)v; }
};
template <> struct St<f16s> {
    template <class C> __device__ static C ld(const f16s* p, long i) { return (C)f16_to_float(p[i].b); }
    template <class C> __device__ static void st(f16s* p, long i, C v) { p[i].b = float_to_f16((float)v); }
};
template <> struct St<bf16s> {
    template <class C> __device__ static C ld(const bf16s* p, long i) { return (C)bf16_to_float(p[i].b); }
    template <class C> __device__ static void st(bf16s* p, long i, C v) { p[i].b = float_to_bf16((float)v); }
};

// ------------------------------------------------------------------ binary16 arithmetic
// Native-half FSAI setup is modelled as IEEE binary16 arithmetic: each + - * / sqrt
// is computed in float and rounded to binary16 (round to nearest even).  float has
// 24 >= 2*11+2 significand bits, so this equals correctly rounded half arithmetic.
struct hnum {
    float v;
    __host__ __device__ hnum() : v(0.0f) {}
    __device__ explicit hnum(float x) : v(f16_to_float(float_to_f16(x))) {}
    __device__ explicit hnum(int x) : v(f16_to_float(float_to_f16((float)x))) {}
    __host__ __device__ static hnum raw(float f) { hnum h; h.v = f; return h; }
    __device__ explicit operator float() const { return v; }
    __device__ explicit operator double() const { return v; }
};
__device__ __forceinline__ hnum operator+(hnum a, hnum b) { return hnum(a.v + b.v); }
__device__ __forceinline__ hnum operator-(hnum a, hnum b) { return hnum(a.v - b.v); }
__device__ __forceinline__ hnum operator*(hnum a, hnum b) { return hnum(a.v * b.v); }
__device__ __forceinline__ hnum operator/(hnum a, hnum b) { return hnum(a.v / b.v); }
__device__ __forceinline__ hnum& operator+=(hnum& a, hnum b) { a = a + b; return a; }
__device__ __forceinline__ hnum& operator-=(hnum& a, hnum b) { a = a - b; return a; }
__device__ __forceinline__ bool operator>(hnum a, hnum b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<(hnum a, hnum b) { return a.v < b.v; }
__device__ __forceinline__ bool operator<=(hnum a, hnum b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator==(hnum a, hnum b) { return a.v == b.v; }
__device__ __forceinline__ bool operator!=(hnum a, hnum b) { return a.v != b.v; }

__device__ __forceinline__ double nsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float nsqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ hnum nsqrt(hnum a) { return hnum(sqrtf(a.v)); }
__device__ __forceinline__ double nabs(double a) { return fabs(a); }
__device__ __forceinline__ float nabs(float a) { return fabsf(a); }
__device__ __forceinline__ hnum nabs(hnum a) { return hnum::raw(fabsf(a.v)); }
__device__ __forceinline__ bool nfinite(double a) { return isfinite(a); }
__device__ __forceinline__ bool nfinite(float a) { return isfinite(a); }
__device__ __forceinline__ bool nfinite(hnum a) { return isfinite(a.v); }

// Storage type -> setup arithmetic type.
template <class S> struct Arith { using type = S; };
template <> struct Arith<f16s> { using type = hnum; };

template <class C, class S> __device__ __forceinline__ C ldv(const S* p, long i);
template <> __device__ __forceinline__ double ldv<double, double>(const double* p, long i) { return p[i]; }
template <> __device__ __forceinline__ float ldv<float, float>(const float* p, long i) { return p[i]; }
template <> __device__ __forceinline__ hnum ldv<hnum, f16s>(const f16s* p, long i) { return hnum::raw(f16_to_float(p[i].b)); }
template <class S, class C> __device__ __forceinline__ void stv(S* p, long i, C v);
template <> __device__ __forceinline__ void stv<double, double>(double* p, long i, double v) { p[i] = v; }
template <> __device__ __forceinline__ void stv<float, float>(float* p, long i, float v) { p[i] = v; }
template <> __device__ __forceinline__ void stv<float, double>(float* p, long i, double v) { p[i] = (float)v; }
template <> __device__ __forceinline__ void stv<f16s, hnum>(f16s* p, long i, hnum v) { p[i].b = float_to_f16(v.v); }

// ------------------------------------------------------------------ wave primitives
// Orders LDS traffic between the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float shfl(float v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ double shfl(double v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ int shfl(int v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ hnum shfl(hnum v, int src) { return hnum::raw(__shfl(v.v, src, kWave)); }

__device__ __forceinline__ float shfl_xor(float v, int m, int w) { return __shfl_xor(v, m, w); }
__device__ __forceinline__ double shfl_xor(double v, int m, int w) { return __shfl_xor(v, m, w); }

// ------------------------------------------------------------------ launch plumbing
// Gathers through a buffer descriptor (sliced products): 32-bit byte offsets from
// the descriptor's base; an offset at or past its range returns 0 and touches no
// memory, which is how padding entries contribute nothing.
template <class TX> __device__ __forceinline__ TX buffer_ld(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ __forceinline__ double buffer_ld<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <> __device__ __forceinline__ float buffer_ld<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
// Word 3 of a gfx9-family buffer descriptor (32-bit data format, no swizzle).
constexpr int kBufferWord3 = 0x00020000;
// Descriptor range of int32 columns: a real column's offset is below 2^31 (the
// host keeps ncol * row stride <= 2^31), a padding column's (-1) is not.
constexpr unsigned kRange32 = 0x80000000u;

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int launch_status() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? WPC_OK : (int)e;
}

inline unsigned blocks_for(long work, int per_block) {
    long b = (work + per_block - 1) / per_block;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace wpc
