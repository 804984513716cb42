// afdtd_common.cuh — device helpers shared by the stencil kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "afdtd_internal.h"

// Storage element types.  Arithmetic is always FP32; conversion to 16-bit
// storage rounds to nearest even.
typedef __hip_bfloat16 bf16_t;
typedef _Float16 f16_t;

template <typename T> struct Store;
// Raw / ldr / cvt split a load from its conversion, so a value loaded one
// loop iteration ahead stays in flight until it is used.
template <> struct Store<float> {
    typedef float Raw;
    static __device__ __forceinline__ Raw ldr(const float* p, int i) { return p[i]; }
    static __device__ __forceinline__ float cvt(Raw r) { return r; }
    static __device__ __forceinline__ Raw raw_of(float v) { return v; }
    static __device__ __forceinline__ float ld(const float* p, int i) { return p[i]; }
    static __device__ __forceinline__ void st(float* p, int i, float v) { p[i] = v; }
    // streaming store: written once per step, not re-read by this sweep
    static __device__ __forceinline__ void st_nt(float* p, int i, float v) {
        __builtin_nontemporal_store(v, p + i);
    }
    static constexpr bool exact = true;
};
template <> struct Store<bf16_t> {
    typedef uint16_t Raw;
    static __device__ __forceinline__ Raw ldr(const bf16_t* p, int i) {
        return ((const uint16_t*)p)[i];
    }
    static __device__ __forceinline__ float cvt(Raw r) {
        return __uint_as_float((uint32_t)r << 16);
    }
    // the bits st() would store (round to nearest even)
    static __device__ __forceinline__ Raw raw_of(float v) {
        return __bfloat16_as_ushort(__float2bfloat16(v));
    }
    static __device__ __forceinline__ float ld(const bf16_t* p, int i) {
        return __bfloat162float(p[i]);
    }
    static __device__ __forceinline__ void st(bf16_t* p, int i, float v) {
        p[i] = __float2bfloat16(v);
    }
    static __device__ __forceinline__ void st_nt(bf16_t* p, int i, float v) { st(p, i, v); }
    static constexpr bool exact = false;
};
template <> struct Store<f16_t> {
    typedef f16_t Raw;
    static __device__ __forceinline__ Raw ldr(const f16_t* p, int i) { return p[i]; }
    static __device__ __forceinline__ float cvt(Raw r) { return (float)r; }
    static __device__ __forceinline__ Raw raw_of(float v) { return (f16_t)v; }
    static __device__ __forceinline__ float ld(const f16_t* p, int i) { return (float)p[i]; }
    static __device__ __forceinline__ void st(f16_t* p, int i, float v) { p[i] = (f16_t)v; }
    static __device__ __forceinline__ void st_nt(f16_t* p, int i, float v) { st(p, i, v); }
    static constexpr bool exact = false;
};

// Block floating point (bfp16): a field of N cells is N int16 mantissas
// followed (at the next 4-byte boundary) by one int32 exponent per 64-cell
// z-row segment (row index (i*Ny + j)*nrz + k/64, nrz = ceil(Nz/64)); a value
// is q * 2^e.  A segment is written as a whole: e is the smallest exponent
// with max|v| * 2^-e < 2^15 and q = rint(v * 2^-e), so every value carries an
// absolute error of at most 2^-16 of its segment's maximum.
struct bfp16_t {
    int16_t q;
};
template <> struct Store<bfp16_t> {
    typedef int16_t Raw;
    static __device__ __forceinline__ Raw ldr(const bfp16_t* p, int i) {
        return ((const int16_t*)p)[i];
    }
    static constexpr bool exact = false;
};
// FP16 with a per-row power-of-two scale (f16s): the layout and exponent
// rule of bfp16 with FP16 values in place of the int16 mantissas; a value is
// h * 2^e, h = v * 2^-e rounded to nearest even in FP16.  The exponent puts
// the segment's largest magnitude in [2^14, 2^15), below FP16's largest
// finite value, and values down to 2^-28 of it stay normal (relative error
// <= 2^-11 each).
struct f16s_t {
    uint16_t h;
};
template <> struct Store<f16s_t> {
    typedef f16_t Raw;
    static __device__ __forceinline__ Raw ldr(const f16s_t* p, int i) {
        return ((const f16_t*)p)[i];
    }
    static constexpr bool exact = false;
};
template <typename T> struct IsBlock { static constexpr bool value = false; };
template <> struct IsBlock<bfp16_t> { static constexpr bool value = true; };
template <> struct IsBlock<f16s_t> { static constexpr bool value = true; };

__host__ __device__ __forceinline__ size_t bfp_exp_offset(size_t N) { return (N * 2 + 3) & ~(size_t)3; }
__device__ __forceinline__ float bfp_dec(int16_t q, int e) { return ldexpf((float)q, e); }
// exponent of a segment whose largest magnitude is m (m >= 0, or NaN / Inf,
// which give the largest exponent so that the stability check sees them)
__device__ __forceinline__ int bfp_exp(float m) {
    const int biased = (int)((__float_as_uint(m) >> 23) & 0xFF);
    return max(biased - 141, -128);
}
__device__ __forceinline__ int16_t bfp_enc(float v, int e) {
    const float x = rintf(ldexpf(v, -e));
    return (int16_t)fminf(fmaxf(x, -32767.0f), 32767.0f);
}
// the same pair for f16s (NaN and infinities pass through, so the stability
// check sees them)
__device__ __forceinline__ float bfp_dec(f16_t h, int e) { return ldexpf((float)h, e); }
__device__ __forceinline__ f16_t f16s_enc(float v, int e) { return (f16_t)ldexpf(v, -e); }
// encode for the block type whose stored word is Raw
template <typename Raw> __device__ __forceinline__ Raw blk_enc(float v, int e);
template <> __device__ __forceinline__ int16_t blk_enc<int16_t>(float v, int e) { return bfp_enc(v, e); }
template <> __device__ __forceinline__ f16_t blk_enc<f16_t>(float v, int e) { return f16s_enc(v, e); }
// Largest |v| over the 64 lanes of the wavefront (all lanes must be active;
// the result is wave-uniform).  DPP row shifts 1, 2, 4, 8 leave the maximum
// of each 16-lane row in its lane 15 (lanes without a source keep 0, the
// identity); the four rows are then combined through readlane.
__device__ __forceinline__ float wave_absmax(float v) {
    int m = (int)(__float_as_uint(v) & 0x7FFFFFFFu);      // |v| bits: order of magnitudes
    m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x111, 0xF, 0xF, false));   // row_shr:1
    m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x112, 0xF, 0xF, false));   // row_shr:2
    m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x114, 0xF, 0xF, false));   // row_shr:4
    m = max(m, __builtin_amdgcn_update_dpp(0, m, 0x118, 0xF, 0xF, false));   // row_shr:8
    const int r = max(max(__builtin_amdgcn_readlane(m, 15), __builtin_amdgcn_readlane(m, 31)),
                      max(__builtin_amdgcn_readlane(m, 47), __builtin_amdgcn_readlane(m, 63)));
    return __uint_as_float((unsigned)r);
}
// Exponent a single value would need on its own (the segment exponent is
// the largest of these over the segment's lanes).
__device__ __forceinline__ int bfp_lane_exp(float v) { return bfp_exp(fabsf(v)); }
// Segment exponent of the wavefront's 64 values, given the exponent the same
// segment had before (e_old, wave-uniform): when no valid lane needs more
// and one needs exactly e_old, the result is e_old (two ballots); otherwise
// the full reduction.  Equal to bfp_exp(wave_absmax(...)) in every case.
__device__ __forceinline__ int bfp_row_exp(float v, bool valid, int e_old) {
    const int le = valid ? bfp_lane_exp(v) : -1000;
    if (!__any(le > e_old) && __any(le == e_old)) return e_old;
    return bfp_exp(wave_absmax(valid ? v : 0.0f));
}
__device__ __forceinline__ const int* bfp_exps(const void* F, int N) {
    return (const int*)((const char*)F + bfp_exp_offset((size_t)N));
}
__device__ __forceinline__ int* bfp_exps(void* F, int N) {
    return (int*)((char*)F + bfp_exp_offset((size_t)N));
}
// element c of a block field (aux kernels; rows are found from c)
template <typename T>
__device__ __forceinline__ float bfp_ld(const void* F, int c, int N, int Nz, int nrz) {
    typedef typename Store<T>::Raw Raw;
    const int row = (c / Nz) * nrz + (c % Nz) / 64;
    return bfp_dec(((const Raw*)F)[c], bfp_exps(F, N)[row]);
}

// Cell geometry for element access to block-floating-point fields.
struct FieldGeo {
    int N, Nz, nrz;
};
inline FieldGeo field_geo(int N, int Nz) { return FieldGeo{N, Nz, (Nz + 63) / 64}; }
// element c of a field of any storage type (FP32 value)
template <typename T>
__device__ __forceinline__ float fld(const T* F, int c, FieldGeo) { return Store<T>::ld(F, c); }
template <>
__device__ __forceinline__ float fld<bfp16_t>(const bfp16_t* F, int c, FieldGeo g) {
    return bfp_ld<bfp16_t>(F, c, g.N, g.Nz, g.nrz);
}
template <>
__device__ __forceinline__ float fld<f16s_t>(const f16s_t* F, int c, FieldGeo g) {
    return bfp_ld<f16s_t>(F, c, g.N, g.Nz, g.nrz);
}

__device__ __forceinline__ int is_solid(uint8_t c) { return c == 1 || c == 2; }

// PML depth of one axis position (acoustic_fdtd.cpp:381-383).
__device__ __forceinline__ int pml_depth(int i, int n, int np) {
    return (i < np) ? (np - i) : ((i >= n - np) ? (i - (n - np) + 1) : 0);
}
__device__ __forceinline__ int imax3(int a, int b, int c) {
    int d = a > b ? a : b;
    return d > c ? d : c;
}

// Face damping 1 - 0.5*(aL + aR)*dt (acoustic_fdtd.cpp:542).
__device__ __forceinline__ float face_damp(const float* alpha, int dL, int dR, float dt) {
    return 1.0f - 0.5f * (alpha[dL] + alpha[dR]) * dt;
}

// Locally reacting wall update of a boundary cell (order 2, extension):
// P' = ((P (1 - xi)) - kap*dv) / (1 + xi), xi = sum over the cell's domain
// sides of (kap/2) * g_side / sqrt(rho K), g = (1 - R)/(1 + R).
__device__ __forceinline__ float impedance_update(float P, float dv, float kap, float xi) {
    float num = P * (1.0f - xi) - kap * dv;
    return num / (1.0f + xi);
}

// Sum of the homogeneous wall terms of the sides a cell touches, accumulated
// in the fixed side order x-, x+, y-, y+, z-, z+.
__device__ __forceinline__ float wall_xi(const float* t, int i, int j, int k, int Nx, int Ny,
                                         int Nz) {
    float xi = 0.0f;
    if (i == 0) xi = xi + t[0];
    if (i == Nx - 1) xi = xi + t[1];
    if (j == 0) xi = xi + t[2];
    if (j == Ny - 1) xi = xi + t[3];
    if (k == 0) xi = xi + t[4];
    if (k == Nz - 1) xi = xi + t[5];
    return xi;
}
