// afdtd_aux.hip — the small kernels around the stencil: Kirchhoff plate,
// plate-velocity injection, point sources + receiver gather, the sampled
// stability check, API gathers, storage conversions and the FP64 energy audit.
#include "afdtd_common.cuh"

#include <vector>

namespace {

// Kirchhoff plate leapfrog over the active nodes (acoustic_fdtd.cpp:984-1053).
// The 13-tap biharmonic uses the precomputed redirected indices
// (acoustic_fdtd.cpp:673-719) and sums in tap order.
__global__ void __launch_bounds__(256)
k_plate_step(int nact, const int* __restrict__ act, const int* __restrict__ b13,
             const int* __restrict__ above, const int* __restrict__ below,
             const float* __restrict__ P, const float* __restrict__ w,
             const float* __restrict__ wp, float* __restrict__ wt,
             const float* __restrict__ ext, float a, float bhc, float vd, float bk,
             int* status, int step_index, float* __restrict__ vpl, float inv_dt) {
    if (status[0]) return;
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nact) return;
    const float W13[13] = {20.f, -8.f, -8.f, -8.f, -8.f, 2.f, 2.f, 2.f, 2.f, 1.f, 1.f, 1.f, 1.f};
    const int p = act[n];
    const int* nb = b13 + 13 * n;
    float bh = 0.0f;
    for (int t = 0; t < 13; ++t) bh += W13[t] * w[nb[t]];
    float Pa = above[n] >= 0 ? P[above[n]] : 0.0f;
    float Pb = below[n] >= 0 ? P[below[n]] : 0.0f;
    float fb = ext[p];
    float bhp = 0.0f;
    for (int t = 0; t < 13; ++t) bhp += W13[t] * wp[nb[t]];
    float vp = w[p] - wp[p];
    float wn = w[p] + vd * vp + a * ((Pb - Pa) + fb) - bhc * bh - bk * (bh - bhp);
    if (!isfinite(wn) || fabsf(wn) > 0.02f) {
        // Plate guard (acoustic_fdtd.cpp:1032-1040): the step fails before the
        // pressure update; later kernels see the flag and skip.
        if (atomicCAS(&status[0], 0, 2) == 0) status[1] = step_index;
        return;
    }
    wt[p] = wn;
    // plate velocity of this step, as the face injection computes it from the
    // rotated buffers: (w_new - w) / dt
    if (vpl) vpl[p] = (wn - w[p]) * inv_dt;
}

// Code byte of every cell for the single-step sweep with cell geometry (bit
// layout in afdtd_fused_tile.cuh): wall / plate cell, closed faces (fraction
// == 0, the closed-face rule of the generic sweep) and faces whose
// divergence tap must be multiplied by the fraction.  The multiply is skipped
// only where it cannot change a bit: fraction exactly 1, or exactly +0 on a
// closed face (whose velocity is +0 after the update).
__global__ void __launch_bounds__(256)
k_geo_codes(int Nx, int Ny, int Nz, const uint8_t* __restrict__ ct,
            const float* __restrict__ fx, const float* __restrict__ fy,
            const float* __restrict__ fz, int has_closed, int has_frac,
            uint8_t* __restrict__ code) {
    const size_t N = (size_t)Nx * Ny * Nz;
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const int k = (int)(c % Nz);
    const size_t t = c / Nz;
    const int j = (int)(t % Ny), i = (int)(t / Ny);
    uint32_t g = 0;
    if (ct && is_solid(ct[c])) g |= 1u;
    const bool exists[3] = {i < Nx - 1, j < Ny - 1, k < Nz - 1};
    const float* F[3] = {fx, fy, fz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!exists[a] || !F[a]) continue;
        const float f = F[a][c];
        const uint32_t b = __float_as_uint(f);
        const bool closed = has_closed && f == 0.0f;
        if (closed) g |= 2u << a;
        if (has_frac && b != 0x3F800000u && !(closed && b == 0u)) g |= 16u << a;
    }
    code[c] = (uint8_t)g;
}

// Plate faces (the Vz faces the plate velocity overrides): plate bit, and the
// fraction multiply unless the fraction is exactly 1 (the plate velocity
// replaces the closed face's 0).
__global__ void __launch_bounds__(256)
k_geo_plate(int nact, const int* __restrict__ vzb, const int* __restrict__ vza,
            const float* __restrict__ fz, int has_frac, uint8_t* __restrict__ code) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nact) return;
    const int f[2] = {vzb[n], vza[n]};
    for (int s = 0; s < 2; ++s) {
        if (f[s] < 0) continue;
        uint32_t g = code[f[s]] | 128u;
        if (has_frac && __float_as_uint(fz[f[s]]) != 0x3F800000u) g |= 64u;
        code[f[s]] = (uint8_t)g;
    }
}

// Plate velocity as the normal velocity of the two plate-layer Vz faces
// (acoustic_fdtd.cpp:1158-1163).  vzb/vza are padded-layout indices or -1.
__global__ void __launch_bounds__(256)
k_plate_inject(int nact, const int* __restrict__ act, const int* __restrict__ vzb,
               const int* __restrict__ vza, const float* __restrict__ w,
               const float* __restrict__ wp, float inv_dt, float* __restrict__ Vz,
               const int* status) {
    if (status[0]) return;
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nact) return;
    int p = act[n];
    float v = (w[p] - wp[p]) * inv_dt;
    if (vzb[n] >= 0) Vz[vzb[n]] = v;
    if (vza[n] >= 0) Vz[vza[n]] = v;
}

// After the pressure update of step `step_index`: point sources (sequential,
// additive) then the 8-tap receiver gather in the summation order of
// fdtd_sample_pressure_precomputed (acoustic_fdtd.cpp:1672-1673).
template <typename T>
__global__ void __launch_bounds__(256)
k_post_step(T* __restrict__ P, int npsrc, const int* __restrict__ pidx,
            const float* __restrict__ psig, int plen, int nrec,
            const int* __restrict__ ridx, const float* __restrict__ rwgt,
            float* __restrict__ trace, int trace_row, int step_index, const int* status,
            FieldGeo g) {
    typedef Store<T> S;
    if (status[0]) return;
    if constexpr (IsBlock<T>::value) {
        // Block storage: the first wavefront re-encodes the 64-cell row of
        // each source cell after the addition, one source after another.
        if (threadIdx.x < 64 && step_index < plen) {
            const int lane = threadIdx.x;
            typedef typename S::Raw Raw;
            Raw* q = (Raw*)P;
            int* E = bfp_exps((void*)P, g.N);
            for (int s = 0; s < npsrc; ++s) {
                const int c = pidx[s];
                const int line = c / g.Nz, kk = c % g.Nz, seg = kk / 64;
                const int k = seg * 64 + lane;
                const bool valid = k < g.Nz;
                const int cell = line * g.Nz + k, row = line * g.nrz + seg;
                float v = valid ? bfp_dec(q[cell], E[row]) : 0.0f;
                if (k == kk) v = v + psig[(size_t)s * plen + step_index];
                const int e = bfp_exp(wave_absmax(v));
                if (valid) q[cell] = blk_enc<Raw>(v, e);
                if (lane == 0) E[row] = e;
                __threadfence();      // the next source may re-read this row
            }
        }
    } else {
        if (threadIdx.x == 0 && step_index < plen) {
            for (int s = 0; s < npsrc; ++s) {
                int c = pidx[s];
                S::st(P, c, S::ld(P, c) + psig[(size_t)s * plen + step_index]);
            }
        }
    }
    __threadfence_block();
    __syncthreads();
    if (trace == nullptr || trace_row < 0) return;
    for (int r = threadIdx.x; r < nrec; r += blockDim.x) {
        const int* id = ridx + 8 * r;
        const float* w = rwgt + 8 * r;
        float v = w[0] * fld(P, id[0], g) + w[1] * fld(P, id[1], g) + w[2] * fld(P, id[2], g)
                + w[3] * fld(P, id[3], g) + w[4] * fld(P, id[4], g) + w[5] * fld(P, id[5], g)
                + w[6] * fld(P, id[6], g) + w[7] * fld(P, id[7], g);
        trace[(size_t)trace_row * nrec + r] = v;
    }
}

// Sampled stability check (acoustic_fdtd.cpp:1245-1268): cells start,
// start+1013, ...; non-finite or |p| > 2e3 Pa fails the step.
template <typename T>
__global__ void __launch_bounds__(256)
k_stability(const T* __restrict__ P, int N, int start, int* status, int step_index,
            FieldGeo g) {
    if (status[0]) return;
    long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long q = start + m * 1013;
    if (q >= N) return;
    float p = fld(P, (int)q, g);
    if (!isfinite(p) || p > 2e3f || p < -2e3f) {
        if (atomicCAS(&status[0], 0, 3) == 0) status[1] = step_index;
    }
}

__global__ void k_fill(float* p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Force injection ext[p] += (scale * val[i]) * (1/dx^2) (acoustic_fdtd.cpp:
// 1716-1719, :1735-1740), grouped by target node: node grp[ng + 1 + g]
// receives its entries grp[2 ng + 1 + q], q in [grp[g], grp[g+1]), in call
// order, so every node's sum is the reference's sequential loop's and the
// nodes run in parallel.
__global__ void k_accumulate_groups(float* ext, const int* __restrict__ grp, int ng,
                                    const float* __restrict__ val, float scale, float inv_dx2,
                                    const int* status) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (status[0] || g >= ng) return;
    const int* start = grp;
    const int* node = grp + ng + 1;
    const int* ent = grp + 2 * ng + 1;
    const int p = node[g];
    float acc = ext[p];
    for (int q = start[g]; q < start[g + 1]; ++q) acc += scale * val[ent[q]] * inv_dx2;
    ext[p] = acc;
}

template <typename T>
__global__ void k_gather(const T* __restrict__ F, const int* __restrict__ idx,
                         const float* __restrict__ wgt, int n, int taps, int zero_start,
                         float scale, float* __restrict__ out, FieldGeo g) {
    typedef Store<T> S;
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int* id = idx + taps * r;
    const float* w = wgt + taps * r;
    auto ld = [&](int c) { return S::exact ? fld(F, c, g) : fld(F, c, g) * scale; };
    float v;
    if (zero_start) {
        v = 0.0f;
        for (int t = 0; t < taps; ++t) v += w[t] * ld(id[t]);
    } else {
        v = w[0] * ld(id[0]);
        for (int t = 1; t < taps; ++t) v = v + w[t] * ld(id[t]);
    }
    out[r] = v;
}

template <typename T>
__global__ void k_to_f32(const T* __restrict__ src, float* __restrict__ dst, size_t n, float scale,
                         FieldGeo g) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = fld(src, (int)i, g) * scale;
}

// FP32 field -> block storage, one wavefront per 64-cell row segment.
template <typename T>
__global__ void __launch_bounds__(256)
k_from_f32_bfp(const float* __restrict__ src, T* __restrict__ dst, int nrows, FieldGeo g) {
    typedef typename Store<T>::Raw Raw;
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrows) return;                   // whole wavefront leaves together
    const int line = row / g.nrz, seg = row - line * g.nrz;
    const int k = seg * 64 + lane;
    const bool valid = k < g.Nz;
    const int cell = line * g.Nz + k;
    const float v = valid ? src[cell] : 0.0f;
    const int e = bfp_exp(wave_absmax(v));
    if (valid) ((Raw*)dst)[cell] = blk_enc<Raw>(v, e);
    if (lane == 0) bfp_exps((void*)dst, g.N)[row] = e;
}

template <typename T>
__global__ void k_from_f32(const float* __restrict__ src, T* __restrict__ dst, size_t n, float scale) {
    typedef Store<T> S;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) S::st(dst, (int)i, src[i] * scale);
}

// FP64 energy audit: out[0] += sum p^2/(2K) dV (non-solid cells),
// out[1] += sum rho_f v^2/2 dV (valid faces).  Block reduction in double.
template <typename T>
__global__ void __launch_bounds__(256)
k_energy(StencilParams p, StencilAux a, const T* __restrict__ P, const T* __restrict__ Vx,
         const T* __restrict__ Vy, const T* __restrict__ Vz, double dV, double* out, FieldGeo g) {
    typedef Store<T> S;
    __shared__ double red[2][256];
    const int tid = threadIdx.x;
    const int N = p.Nx * p.Ny * p.Nz;
    double ep = 0.0, ek = 0.0;
    for (int c = blockIdx.x * blockDim.x + tid; c < N; c += gridDim.x * blockDim.x) {
        int k = c % p.Nz, t = c / p.Nz, j = t % p.Ny, i = t / p.Ny;
        if (!(p.has_solid && is_solid(a.ct[c]))) {
            double K = p.hetero ? (double)a.kap[c] / ((double)p.dt / (double)p.dx) : (double)p.K0;
            double pc = (double)fld(P, c, g);
            ep += pc * pc / (2.0 * K);
        }
        double vs = S::exact ? 1.0 : (double)p.inv_sV;
        // face density = 1 / face buoyancy (exact energy weight of the scheme)
        auto rf = [&](int L, int R) {
            return p.hetero ? 2.0 / ((double)a.bu[L] + (double)a.bu[R]) : (double)p.rho_air;
        };
        if (i < p.Nx - 1) {
            double v = (double)fld(Vx, c, g) * vs;
            ek += 0.5 * rf(c, c + p.Ny * p.Nz) * v * v;
        }
        if (j < p.Ny - 1) {
            double v = (double)fld(Vy, c, g) * vs;
            ek += 0.5 * rf(c, c + p.Nz) * v * v;
        }
        if (k < p.Nz - 1) {
            double v = (double)fld(Vz, c, g) * vs;
            ek += 0.5 * rf(c, c + 1) * v * v;
        }
    }
    red[0][tid] = ep; red[1][tid] = ek;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) { red[0][tid] += red[0][tid + s]; red[1][tid] += red[1][tid + s]; }
        __syncthreads();
    }
    if (tid == 0) {
        atomicAdd(&out[0], red[0][0] * dV);
        atomicAdd(&out[1], red[1][0] * dV);
    }
}

inline unsigned blocks_for(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t launch_plate_step(AcousticFDTDState* st, int step_index) {
    if (st->nact <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plate_step, dim3(blocks_for(st->nact, 256)), dim3(256), 0, st->stream,
                       st->nact, st->d_act, st->d_b13, st->d_above, st->d_below,
                       (const float*)st->P[st->cur], st->d_w[0], st->d_w[1], st->d_w[2],
                       st->d_ext, st->plate_a, st->plate_bh, st->plate_vd, st->plate_bk,
                       st->d_status, step_index, st->fused_geo ? st->d_vpl : nullptr,
                       1.0f / st->dt);
    return hipGetLastError();
}

hipError_t launch_geo_codes(AcousticFDTDState* st) {
    if (!st->d_geo) {
        hipError_t e = hipMalloc(&st->d_geo, (size_t)st->N);
        if (e != hipSuccess) { st->d_geo = nullptr; return e; }
    }
    hipLaunchKernelGGL(k_geo_codes, dim3(blocks_for((size_t)st->N, 256)), dim3(256), 0,
                       st->stream, st->Nx, st->Ny, st->Nz, st->has_solid ? st->d_ct : nullptr,
                       st->d_fx, st->d_fy, st->d_fz, st->sp.has_closed, st->has_frac, st->d_geo);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || st->nact <= 0) return e;
    if (!st->d_vpl) {
        e = hipMalloc(&st->d_vpl, sizeof(float) * st->Nplate);
        if (e != hipSuccess) { st->d_vpl = nullptr; return e; }
    }
    hipLaunchKernelGGL(k_geo_plate, dim3(blocks_for(st->nact, 256)), dim3(256), 0, st->stream,
                       st->nact, st->d_vzb, st->d_vza, st->d_fz, st->has_frac, st->d_geo);
    return hipGetLastError();
}

hipError_t launch_plate_inject(AcousticFDTDState* st) {
    if (st->nact <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_plate_inject, dim3(blocks_for(st->nact, 256)), dim3(256), 0,
                       st->stream, st->nact, st->d_act, st->d_vzb, st->d_vza, st->d_w[0],
                       st->d_w[1], 1.0f / st->dt, (float*)st->V[2], st->d_status);
    return hipGetLastError();
}

hipError_t launch_post_step(AcousticFDTDState* st, int step_index, int trace_row) {
    if (st->npsrc == 0 && (st->nrec == 0 || trace_row < 0)) return hipSuccess;
    float* tr = (st->nrec > 0 && trace_row >= 0) ? st->d_trace : nullptr;
    void* P = st->P[st->cur];
    const FieldGeo g = field_geo(st->N, st->Nz);
#define POST(T)                                                                         \
    hipLaunchKernelGGL(k_post_step<T>, dim3(1), dim3(256), 0, st->stream, (T*)P,        \
                       st->npsrc, st->d_pidx, st->d_psig, st->plen, st->nrec, st->d_ridx, \
                       st->d_rwgt, tr, trace_row, step_index, st->d_status, g)
    switch (st->storage) {
        case FDTD_STORE_BF16: POST(bf16_t); break;
        case FDTD_STORE_F16: POST(f16_t); break;
        case FDTD_STORE_BFP16: POST(bfp16_t); break;
        case FDTD_STORE_F16S: POST(f16s_t); break;
        default: POST(float);
    }
#undef POST
    return hipGetLastError();
}

hipError_t launch_stability_check(AcousticFDTDState* st, int start, int step_index) {
    long cnt = (st->N > start) ? ((long)(st->N - start) + 1012) / 1013 : 0;
    if (cnt <= 0) return hipSuccess;
    const void* P = st->P[st->cur];
    const FieldGeo g = field_geo(st->N, st->Nz);
#define CHK(T)                                                                          \
    hipLaunchKernelGGL(k_stability<T>, dim3(blocks_for(cnt, 256)), dim3(256), 0, st->stream, \
                       (const T*)P, st->N, start, st->d_status, step_index, g)
    switch (st->storage) {
        case FDTD_STORE_BF16: CHK(bf16_t); break;
        case FDTD_STORE_F16: CHK(f16_t); break;
        case FDTD_STORE_BFP16: CHK(bfp16_t); break;
        case FDTD_STORE_F16S: CHK(f16s_t); break;
        default: CHK(float);
    }
#undef CHK
    return hipGetLastError();
}

hipError_t launch_fill(float* p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_accumulate_groups(float* ext, const int* grp, int ng, const float* val,
                                    float scale, float inv_dx2, const int* status,
                                    hipStream_t s) {
    if (ng <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_accumulate_groups, dim3(blocks_for(ng, 256)), dim3(256), 0, s, ext, grp,
                       ng, val, scale, inv_dx2, status);
    return hipGetLastError();
}

hipError_t launch_gather(const void* F, int storage, float scale, const int* idx,
                          const float* wgt, int n, int taps, int zero_start, float* out,
                          hipStream_t s, int Ncells, int Nz) {
    if (n <= 0) return hipSuccess;
    const FieldGeo g = field_geo(Ncells, Nz);
#define G(T)                                                                              \
    hipLaunchKernelGGL(k_gather<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const T*)F, \
                       idx, wgt, n, taps, zero_start, scale, out, g)
    switch (storage) {
        case FDTD_STORE_BF16: G(bf16_t); break;
        case FDTD_STORE_F16: G(f16_t); break;
        case FDTD_STORE_BFP16: G(bfp16_t); break;
        case FDTD_STORE_F16S: G(f16s_t); break;
        default: G(float);
    }
#undef G
    return hipGetLastError();
}

hipError_t launch_to_f32(const void* src, int storage, float scale, float* dst, size_t n,
                           hipStream_t s, int Nz) {
    if (n == 0) return hipSuccess;
    const FieldGeo g = field_geo((int)n, Nz);
#define C(T)                                                                               \
    hipLaunchKernelGGL(k_to_f32<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const T*)src, \
                       dst, n, scale, g)
    switch (storage) {
        case FDTD_STORE_BF16: C(bf16_t); break;
        case FDTD_STORE_F16: C(f16_t); break;
        case FDTD_STORE_BFP16: C(bfp16_t); break;
        case FDTD_STORE_F16S: C(f16s_t); break;
        default: C(float);
    }
#undef C
    return hipGetLastError();
}

hipError_t launch_from_f32(const float* src, void* dst, int storage, float scale, size_t n,
                             hipStream_t s, int Nz) {
    if (n == 0) return hipSuccess;
    if (storage == FDTD_STORE_BFP16 || storage == FDTD_STORE_F16S) {
        const FieldGeo g = field_geo((int)n, Nz);
        const int nrows = (int)(n / Nz) * g.nrz;
        if (storage == FDTD_STORE_BFP16)
            hipLaunchKernelGGL(k_from_f32_bfp<bfp16_t>, dim3((nrows + 3) / 4), dim3(256), 0, s,
                               src, (bfp16_t*)dst, nrows, g);
        else
            hipLaunchKernelGGL(k_from_f32_bfp<f16s_t>, dim3((nrows + 3) / 4), dim3(256), 0, s,
                               src, (f16s_t*)dst, nrows, g);
        return hipGetLastError();
    }
#define C(T)                                                                                \
    hipLaunchKernelGGL(k_from_f32<T>, dim3(blocks_for(n, 256)), dim3(256), 0, s, src, (T*)dst, \
                       n, scale)
    switch (storage) {
        case FDTD_STORE_BF16: C(bf16_t); break;
        case FDTD_STORE_F16: C(f16_t); break;
        default: C(float);
    }
#undef C
    return hipGetLastError();
}

hipError_t launch_energy(AcousticFDTDState* st, double* d_out) {
    StencilAux a;
    a.ct = st->d_ct; a.fx = st->d_fx; a.fy = st->d_fy; a.fz = st->d_fz;
    a.bu = st->d_bu; a.kap = st->d_kap; a.xi = st->d_xi;
    a.alpha = st->d_alpha; a.pdamp = st->d_pdamp; a.status = st->d_status;
    double dV = (double)st->dx * (double)st->dx * (double)st->dx;
    unsigned nb = blocks_for(st->N, 256);
    if (nb > 2048) nb = 2048;
    const FieldGeo g = field_geo(st->N, st->Nz);
#define E(T)                                                                                 \
    hipLaunchKernelGGL(k_energy<T>, dim3(nb), dim3(256), 0, st->stream, st->sp, a,           \
                       (const T*)st->P[st->cur], (const T*)st->V[0], (const T*)st->V[1],     \
                       (const T*)st->V[2], dV, d_out, g)
    switch (st->storage) {
        case FDTD_STORE_BF16: E(bf16_t); break;
        case FDTD_STORE_F16: E(f16_t); break;
        case FDTD_STORE_BFP16: E(bfp16_t); break;
        case FDTD_STORE_F16S: E(f16s_t); break;
        default: E(float);
    }
#undef E
    return hipGetLastError();
}

namespace {
__global__ void __launch_bounds__(256) k_copy16(const float4* __restrict__ a, float4* __restrict__ b,
                                                size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}
struct StreamArrays {
    const float* in[8];
    float* out[8];
};

// R FP32 arrays in, W out, one element per lane and one lane per element:
// the byte mix of a stencil step without the stencil.  R and W are compile-
// time constants and the grid covers the arrays once (no grid-stride loop),
// so the pass is limited by the memory system, not by issue.
template <int R, int W>
__global__ void __launch_bounds__(256) k_stream_mix(StreamArrays a, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < R; ++c) s += a.in[c][i];
#pragma unroll
    for (int c = 0; c < W; ++c) __builtin_nontemporal_store(s + (float)c, a.out[c] + i);
}

template <int R>
void launch_mix_w(StreamArrays a, int n_write, size_t n, unsigned blocks, hipStream_t s) {
    switch (n_write) {
        case 1: hipLaunchKernelGGL((k_stream_mix<R, 1>), dim3(blocks), dim3(256), 0, s, a, n); break;
        case 2: hipLaunchKernelGGL((k_stream_mix<R, 2>), dim3(blocks), dim3(256), 0, s, a, n); break;
        case 3: hipLaunchKernelGGL((k_stream_mix<R, 3>), dim3(blocks), dim3(256), 0, s, a, n); break;
        case 4: hipLaunchKernelGGL((k_stream_mix<R, 4>), dim3(blocks), dim3(256), 0, s, a, n); break;
        case 5: hipLaunchKernelGGL((k_stream_mix<R, 5>), dim3(blocks), dim3(256), 0, s, a, n); break;
        case 6: hipLaunchKernelGGL((k_stream_mix<R, 6>), dim3(blocks), dim3(256), 0, s, a, n); break;
        case 7: hipLaunchKernelGGL((k_stream_mix<R, 7>), dim3(blocks), dim3(256), 0, s, a, n); break;
        default: hipLaunchKernelGGL((k_stream_mix<R, 8>), dim3(blocks), dim3(256), 0, s, a, n); break;
    }
}

void launch_mix(StreamArrays a, int n_read, int n_write, size_t n, hipStream_t s) {
    const unsigned blocks = (unsigned)((n + 255) / 256);
    switch (n_read) {
        case 1: launch_mix_w<1>(a, n_write, n, blocks, s); break;
        case 2: launch_mix_w<2>(a, n_write, n, blocks, s); break;
        case 3: launch_mix_w<3>(a, n_write, n, blocks, s); break;
        case 4: launch_mix_w<4>(a, n_write, n, blocks, s); break;
        case 5: launch_mix_w<5>(a, n_write, n, blocks, s); break;
        case 6: launch_mix_w<6>(a, n_write, n, blocks, s); break;
        case 7: launch_mix_w<7>(a, n_write, n, blocks, s); break;
        default: launch_mix_w<8>(a, n_write, n, blocks, s); break;
    }
}
}  // namespace

// Streaming ceiling of a stencil's byte mix (no reference counterpart): n_read
// arrays read and n_write arrays written once per pass over n_cells FP32
// elements, timed with HIP events over `reps` passes.
extern "C" SK_API int fdtd_stream_calibration(size_t n_cells, int n_read, int n_write, int reps,
                                              float* ms_per_pass) {
    if (!ms_per_pass) return FDTD_ERR_NULL;
    if (reps <= 0 || n_cells == 0 || n_read < 1 || n_read > 8 || n_write < 1 || n_write > 8)
        return FDTD_ERR_ARG;
    StreamArrays a{};
    std::vector<void*> bufs;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = FDTD_OK;
    auto fail = [&](hipError_t e) {
        afdtd_set_error(std::string("fdtd_stream_calibration: ") + hipGetErrorString(e));
        rc = FDTD_ERR_DEVICE;
    };
    for (int c = 0; c < n_read + n_write && rc == FDTD_OK; ++c) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, n_cells * sizeof(float));
        if (e != hipSuccess) { fail(e); break; }
        bufs.push_back(p);
        if (c < n_read) a.in[c] = (const float*)p; else a.out[c - n_read] = (float*)p;
    }
    hipError_t e = hipSuccess;
    if (rc == FDTD_OK && (e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) fail(e);
    if (rc == FDTD_OK && (e = hipEventCreate(&e0)) != hipSuccess) fail(e);
    if (rc == FDTD_OK && (e = hipEventCreate(&e1)) != hipSuccess) fail(e);
    for (int c = 0; c < n_read && rc == FDTD_OK; ++c)
        if ((e = hipMemsetAsync((void*)a.in[c], 0, n_cells * sizeof(float), s)) != hipSuccess) fail(e);
    if (rc == FDTD_OK) {
        launch_mix(a, n_read, n_write, n_cells, s);
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) launch_mix(a, n_read, n_write, n_cells, s);
        (void)hipEventRecord(e1, s);
        if ((e = hipGetLastError()) != hipSuccess) fail(e);
        else if ((e = hipStreamSynchronize(s)) != hipSuccess) fail(e);
        float ms = 0.0f;
        if (rc == FDTD_OK && (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) fail(e);
        if (rc == FDTD_OK) *ms_per_pass = ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    for (void* p : bufs) (void)hipFree(p);
    return rc;
}

// Device-bandwidth probe: a 16-byte-per-lane copy of nbytes, timed with HIP
// events over `reps` launches (no reference counterpart; used to report the
// measured copy ceiling next to the stencil's roofline fraction).
extern "C" SK_API int fdtd_copy_calibration(size_t nbytes, int reps, float* ms_per_copy) {
    if (!ms_per_copy || reps <= 0 || nbytes < 16) return FDTD_ERR_NULL;
    const size_t n = nbytes / 16;
    float4 *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    AFDTD_CHECK(hipMalloc(&a, n * 16));
    AFDTD_CHECK(hipMalloc(&b, n * 16));
    AFDTD_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    AFDTD_CHECK(hipEventCreate(&e0));
    AFDTD_CHECK(hipEventCreate(&e1));
    AFDTD_CHECK(hipMemsetAsync(a, 0, n * 16, s));
    const unsigned blocks = 256 * 32;
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, a, b, n);
    AFDTD_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, a, b, n);
    AFDTD_CHECK(hipEventRecord(e1, s));
    AFDTD_CHECK(hipStreamSynchronize(s));
    float ms = 0.0f;
    AFDTD_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_copy = ms / reps;
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipStreamDestroy(s);
    (void)hipFree(a); (void)hipFree(b);
    return FDTD_OK;
}
