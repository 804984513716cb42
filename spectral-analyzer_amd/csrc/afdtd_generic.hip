// afdtd_generic.hip — the general-geometry leapfrog: one velocity sweep and
// one pressure sweep per step.
//
// These kernels carry every feature of the reference step
// (csrc/kernels/acoustic_fdtd.cpp:1081-1241): walls and plate cells, closed
// faces, cut-cell face fractions, PML damping, plus the extension terms
// (order-2 stencil, heterogeneous media, locally reacting walls).  Per cell
// they evaluate the reference's float expressions in the same order, and the
// library is built with -ffp-contract=off, so FP32 results match the
// reference arithmetic bit for bit.  The fused plane-marching kernel
// (afdtd_fused.hip) replaces this pair where the geometry allows it.
//
// Launch: block (64, 4) = 4 wavefronts; one wavefront covers 64 consecutive
// z cells of one (i, j) row, so every field access is a coalesced 256 B
// (FP32) transaction.  Grid (ceil(Nz/64), ceil(Ny/4), Nx).
#include "afdtd_common.cuh"

namespace {

constexpr int BZ = 64, BY = 4;

// Gradient at a face between cells L and R = L + s, with outer taps L - s and
// R + s used when `inner` (acoustic_fdtd.cpp:1107-1112).
template <typename T>
__device__ __forceinline__ float face_gradient(const T* __restrict__ P, int L, int s,
                                               int order, int inner, float C1, float C2) {
    typedef Store<T> S;
    float pl = S::ld(P, L), pr = S::ld(P, L + s);
    if (order == 4 && inner)
        return C1 * (pr - pl) + C2 * (S::ld(P, L + 2 * s) - S::ld(P, L - s));
    return pr - pl;
}

template <typename T>
__global__ void __launch_bounds__(BZ * BY)
k_velocity(StencilParams p, StencilAux a, const T* __restrict__ P, T* __restrict__ Vx,
           T* __restrict__ Vy, T* __restrict__ Vz) {
    typedef Store<T> S;
    if (*a.status) return;
    const int k = blockIdx.x * BZ + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= p.Nz || j >= p.Ny) return;
    const int sy = p.Nz, sx = p.Ny * p.Nz;
    const int c = k + p.Nz * (j + p.Ny * i);
    int di = 0, dj = 0, dk = 0;
    if (p.n_pml > 0) {
        di = pml_depth(i, p.Nx, p.n_pml);
        dj = pml_depth(j, p.Ny, p.n_pml);
        dk = pml_depth(k, p.Nz, p.n_pml);
    }
    // x face (i+1/2, j, k)
    if (i < p.Nx - 1) {
        float g = face_gradient(P, c, sx, p.order, i > 0 && i < p.Nx - 2, p.C1, p.C2);
        float coef = p.hetero ? (0.5f * (a.bu[c] + a.bu[c + sx])) * p.dtdx : p.vel_coeff;
        float v = S::ld(Vx, c);
        if (!S::exact) v = v * p.inv_sV;
        float nv = v - coef * g;
        if (p.n_pml > 0) {
            int dL = imax3(di, dj, dk), dR = imax3(pml_depth(i + 1, p.Nx, p.n_pml), dj, dk);
            nv = nv * face_damp(a.alpha, dL, dR, p.dt);
        }
        if (p.has_closed && a.fx[c] == 0.0f) nv = 0.0f;
        if (!S::exact) nv = nv * p.sV;
        S::st(Vx, c, nv);
    }
    // y face (i, j+1/2, k)
    if (j < p.Ny - 1) {
        float g = face_gradient(P, c, sy, p.order, j > 0 && j < p.Ny - 2, p.C1, p.C2);
        float coef = p.hetero ? (0.5f * (a.bu[c] + a.bu[c + sy])) * p.dtdx : p.vel_coeff;
        float v = S::ld(Vy, c);
        if (!S::exact) v = v * p.inv_sV;
        float nv = v - coef * g;
        if (p.n_pml > 0) {
            int dL = imax3(di, dj, dk), dR = imax3(di, pml_depth(j + 1, p.Ny, p.n_pml), dk);
            nv = nv * face_damp(a.alpha, dL, dR, p.dt);
        }
        if (p.has_closed && a.fy[c] == 0.0f) nv = 0.0f;
        if (!S::exact) nv = nv * p.sV;
        S::st(Vy, c, nv);
    }
    // z face (i, j, k+1/2)
    if (k < p.Nz - 1) {
        float g = face_gradient(P, c, 1, p.order, k > 0 && k < p.Nz - 2, p.C1, p.C2);
        float coef = p.hetero ? (0.5f * (a.bu[c] + a.bu[c + 1])) * p.dtdx : p.vel_coeff;
        float v = S::ld(Vz, c);
        if (!S::exact) v = v * p.inv_sV;
        float nv = v - coef * g;
        if (p.n_pml > 0) {
            int dL = imax3(di, dj, dk), dR = imax3(di, dj, pml_depth(k + 1, p.Nz, p.n_pml));
            nv = nv * face_damp(a.alpha, dL, dR, p.dt);
        }
        if (p.has_closed && a.fz[c] == 0.0f) nv = 0.0f;
        if (!S::exact) nv = nv * p.sV;
        S::st(Vz, c, nv);
    }
}

// Velocity tap of the divergence: V at a face index (or 0 outside the domain,
// acoustic_fdtd.cpp:904-917), weighted by the face fraction when has_frac.
template <typename T>
__device__ __forceinline__ float vtap(const T* __restrict__ V, const float* __restrict__ F,
                                      int idx, bool valid, int has_frac, float inv_sV) {
    typedef Store<T> S;
    if (!valid) return 0.0f;
    float v = S::ld(V, idx);
    if (!S::exact) v = v * inv_sV;
    if (has_frac) v = v * F[idx];
    return v;
}

template <typename T>
__global__ void __launch_bounds__(BZ * BY)
k_pressure(StencilParams p, StencilAux a, const T* __restrict__ P, T* __restrict__ Pn,
           const T* __restrict__ Vx, const T* __restrict__ Vy, const T* __restrict__ Vz) {
    typedef Store<T> S;
    if (*a.status) return;
    const int k = blockIdx.x * BZ + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= p.Nz || j >= p.Ny) return;
    const int sy = p.Nz, sx = p.Ny * p.Nz;
    const int c = k + p.Nz * (j + p.Ny * i);
    if (p.has_solid && is_solid(a.ct[c])) return;   // solid cells keep P = 0
    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz, hf = p.has_frac;
    const float isv = p.inv_sV;
    float xp = vtap(Vx, a.fx, c, i < Nx - 1, hf, isv), xm = vtap(Vx, a.fx, c - sx, i > 0, hf, isv);
    float yp = vtap(Vy, a.fy, c, j < Ny - 1, hf, isv), ym = vtap(Vy, a.fy, c - sy, j > 0, hf, isv);
    float zp = vtap(Vz, a.fz, c, k < Nz - 1, hf, isv), zm = vtap(Vz, a.fz, c - 1, k > 0, hf, isv);
    float dv;
    if (p.order == 2) {
        dv = (xp - xm) + (yp - ym) + (zp - zm);
    } else {
        float xpp = vtap(Vx, a.fx, c + sx, i < Nx - 2, hf, isv);
        float xmm = vtap(Vx, a.fx, c - 2 * sx, i > 1, hf, isv);
        float ypp = vtap(Vy, a.fy, c + sy, j < Ny - 2, hf, isv);
        float ymm = vtap(Vy, a.fy, c - 2 * sy, j > 1, hf, isv);
        float zpp = vtap(Vz, a.fz, c + 1, k < Nz - 2, hf, isv);
        float zmm = vtap(Vz, a.fz, c - 2, k > 1, hf, isv);
        dv = p.C1 * ((xp - xm) + (yp - ym) + (zp - zm))
           + p.C2 * ((xpp - xmm) + (ypp - ymm) + (zpp - zmm));
    }
    const float kap = p.hetero ? a.kap[c] : p.bulk_coeff;
    const float pc = S::ld(P, c);
    float out;
    const bool edge = (i == 0) | (i == Nx - 1) | (j == 0) | (j == Ny - 1) | (k == 0) | (k == Nz - 1);
    if (p.impedance && edge) {
        float xi = p.hetero ? a.xi[c] : wall_xi(p.xterm, i, j, k, Nx, Ny, Nz);
        out = impedance_update(pc, dv, kap, xi);
    } else {
        out = pc - kap * dv;
    }
    if (p.n_pml > 0) {
        int d = imax3(pml_depth(i, Nx, p.n_pml), pml_depth(j, Ny, p.n_pml),
                      pml_depth(k, Nz, p.n_pml));
        out = out * a.pdamp[d];
    }
    S::st(Pn, c, out);
}

StencilAux make_aux(const AcousticFDTDState* st) {
    StencilAux a;
    a.ct = st->d_ct;
    a.fx = st->d_fx; a.fy = st->d_fy; a.fz = st->d_fz;
    a.bu = st->d_bu; a.kap = st->d_kap; a.xi = st->d_xi;
    a.alpha = st->d_alpha; a.pdamp = st->d_pdamp;
    a.status = st->d_status;
    return a;
}

template <typename T>
hipError_t vel_t(AcousticFDTDState* st) {
    dim3 blk(BZ, BY), grd((st->Nz + BZ - 1) / BZ, (st->Ny + BY - 1) / BY, st->Nx);
    hipLaunchKernelGGL(k_velocity<T>, grd, blk, 0, st->stream, st->sp, make_aux(st),
                       (const T*)st->P[st->cur], (T*)st->V[0], (T*)st->V[1], (T*)st->V[2]);
    return hipGetLastError();
}

template <typename T>
hipError_t pres_t(AcousticFDTDState* st) {
    dim3 blk(BZ, BY), grd((st->Nz + BZ - 1) / BZ, (st->Ny + BY - 1) / BY, st->Nx);
    hipLaunchKernelGGL(k_pressure<T>, grd, blk, 0, st->stream, st->sp, make_aux(st),
                       (const T*)st->P[st->cur], (T*)st->P[st->cur ^ 1],
                       (const T*)st->V[0], (const T*)st->V[1], (const T*)st->V[2]);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_velocity_generic(AcousticFDTDState* st) {
    switch (st->storage) {
        case FDTD_STORE_BF16: return vel_t<bf16_t>(st);
        case FDTD_STORE_F16: return vel_t<f16_t>(st);
        default: return vel_t<float>(st);
    }
}

hipError_t launch_pressure_generic(AcousticFDTDState* st) {
    switch (st->storage) {
        case FDTD_STORE_BF16: return pres_t<bf16_t>(st);
        case FDTD_STORE_F16: return pres_t<f16_t>(st);
        default: return pres_t<float>(st);
    }
}
