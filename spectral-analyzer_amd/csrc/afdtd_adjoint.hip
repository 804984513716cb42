// afdtd_adjoint.hip — gradients of receiver traces with respect to the
// medium: a hand-written discrete adjoint of the order-2 leapfrog with
// checkpoint-and-recompute (no reference counterpart; the reference has no
// differentiation).
//
// Forward model (per-cell parameters m = log rho, l = log K, eta = boundary
// admittance sum, REAL = float or double):
//   b = exp(-m), K = exp(l), kappa = K dt/dx,
//   xi = (dt/(2 dx)) * eta * sqrt(K b) on domain-edge cells, 0 elsewhere,
//   beta_f = (b_L + b_R)/2 * dt/dx on every face,
//   V^{n+1/2} = V^{n-1/2} - beta (G P^n)            (G P)_f = P_R - P_L
//   P^{n+1}   = a P^n - e (D V^{n+1/2}) + s_n         a = (1-xi)/(1+xi), e = kappa/(1+xi)
//   y_n       = R P^{n+1}                             (8-tap receiver stencils)
// This is the engine's order-2 step for heterogeneous media and impedance
// walls (per-side g summed into eta), written as a P - e D V.
//
// Adjoint of one step (D^T = -G, G^T = -D), with Pb, Vb the adjoint fields:
//   Pb += R^T gy_n                                    (trace gradient)
//   abar += Pb P^n,  ebar -= Pb (D V^{n+1/2})          (accumulated per cell)
//   Vb += G (e Pb)                                     (velocity-like sweep)
//   betabar_f -= Vb_f (G P^n)_f                        (accumulated per face)
//   Pb  = a Pb + D (beta Vb)                           (pressure-like sweep)
// The forward stores checkpoints (P, V) every K steps; the backward recomputes
// each K-step segment into HBM (P^n and V^{n+1/2} of every step) and runs the
// reverse sweeps over it.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "afdtd_internal.h"

namespace {

constexpr int BZ = 64, BY = 4;

struct AdjGeom {
    int Nx, Ny, Nz;
    double dtdx;        // dt/dx
};

template <typename R>
__global__ void __launch_bounds__(BZ * BY)
k_coef(AdjGeom g, const R* __restrict__ m, const R* __restrict__ l, const R* __restrict__ eta,
       R* __restrict__ b, R* __restrict__ kap, R* __restrict__ xi, R* __restrict__ A,
       R* __restrict__ E) {
    const int k = blockIdx.x * BZ + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, i = blockIdx.z;
    if (k >= g.Nz || j >= g.Ny) return;
    const int c = k + g.Nz * (j + g.Ny * i);
    const R dtdx = (R)g.dtdx;
    const R bc = exp(-m[c]);
    const R K = exp(l[c]);
    const R kp = K * dtdx;
    const bool edge = i == 0 || i == g.Nx - 1 || j == 0 || j == g.Ny - 1 || k == 0 || k == g.Nz - 1;
    const R x = edge ? (R)0.5 * dtdx * eta[c] * sqrt(K * bc) : (R)0;
    b[c] = bc;
    kap[c] = kp;
    xi[c] = x;
    A[c] = ((R)1 - x) / ((R)1 + x);
    E[c] = kp / ((R)1 + x);
}

// V_out = V_in - beta (G P) on the faces owned by each cell.
template <typename R>
__global__ void __launch_bounds__(BZ * BY)
k_fvel(AdjGeom g, const R* __restrict__ P, const R* __restrict__ b, const R* Vxi, const R* Vyi,
       const R* Vzi, R* Vxo, R* Vyo, R* Vzo) {
    const int k = blockIdx.x * BZ + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, i = blockIdx.z;
    if (k >= g.Nz || j >= g.Ny) return;
    const int sy = g.Nz, sx = g.Ny * g.Nz;
    const int c = k + g.Nz * (j + g.Ny * i);
    const R h = (R)0.5 * (R)g.dtdx;
    Vxo[c] = i < g.Nx - 1 ? Vxi[c] - h * (b[c] + b[c + sx]) * (P[c + sx] - P[c]) : (R)0;
    Vyo[c] = j < g.Ny - 1 ? Vyi[c] - h * (b[c] + b[c + sy]) * (P[c + sy] - P[c]) : (R)0;
    Vzo[c] = k < g.Nz - 1 ? Vzi[c] - h * (b[c] + b[c + 1]) * (P[c + 1] - P[c]) : (R)0;
}

template <typename R>
__device__ __forceinline__ R div_at(const R* Vx, const R* Vy, const R* Vz, int c, int i, int j,
                                    int k, const AdjGeom& g) {
    const int sy = g.Nz, sx = g.Ny * g.Nz;
    R d = (R)0;
    if (i < g.Nx - 1) d += Vx[c];
    if (i > 0) d -= Vx[c - sx];
    if (j < g.Ny - 1) d += Vy[c];
    if (j > 0) d -= Vy[c - sy];
    if (k < g.Nz - 1) d += Vz[c];
    if (k > 0) d -= Vz[c - 1];
    return d;
}

// P_out = a P_in - e (D V)
template <typename R>
__global__ void __launch_bounds__(BZ * BY)
k_fpres(AdjGeom g, const R* __restrict__ Pin, R* __restrict__ Pout, const R* __restrict__ Vx,
        const R* __restrict__ Vy, const R* __restrict__ Vz, const R* __restrict__ A,
        const R* __restrict__ E) {
    const int k = blockIdx.x * BZ + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, i = blockIdx.z;
    if (k >= g.Nz || j >= g.Ny) return;
    const int c = k + g.Nz * (j + g.Ny * i);
    Pout[c] = A[c] * Pin[c] - E[c] * div_at(Vx, Vy, Vz, c, i, j, k, g);
}

// Sources after the pressure update; receivers record y_n (forward only).
template <typename R>
__global__ void k_fpost(R* P, int nsrc, const int* sidx, const R* sig, int slen, int n, int nrec,
                        const int* ridx, const R* rwgt, R* traces) {
    if (threadIdx.x == 0 && n < slen)
        for (int s = 0; s < nsrc; ++s) P[sidx[s]] += sig[(size_t)s * slen + n];
    __syncthreads();
    if (!traces) return;
    for (int r = threadIdx.x; r < nrec; r += blockDim.x) {
        R v = (R)0;
        for (int t = 0; t < 8; ++t) v += rwgt[8 * r + t] * P[ridx[8 * r + t]];
        traces[(size_t)n * nrec + r] = v;
    }
}

// Pb += R^T gy_n: one thread per distinct tap cell, adding that cell's
// contributions in receiver-then-tap order (the order of a sequential
// loop over receivers and taps, so the sums are deterministic).
template <typename R>
__global__ void __launch_bounds__(256)
k_inject(R* __restrict__ Pb, int ncell, const int* __restrict__ cell,
         const int* __restrict__ start, const int* __restrict__ rec, const R* __restrict__ w,
         const R* __restrict__ gy) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= ncell) return;
    const int c = cell[u];
    R v = Pb[c];
    for (int q = start[u]; q < start[u + 1]; ++q) v += w[q] * gy[rec[q]];
    Pb[c] = v;
}

// Velocity-like adjoint sweep; also the per-cell (abar, ebar) and per-face
// (betabar) accumulations of this step.
template <typename R>
__global__ void __launch_bounds__(BZ * BY)
k_avel(AdjGeom g, const R* __restrict__ Pb, R* __restrict__ Vbx, R* __restrict__ Vby,
       R* __restrict__ Vbz, const R* __restrict__ Pn, const R* __restrict__ Vx,
       const R* __restrict__ Vy, const R* __restrict__ Vz, const R* __restrict__ E,
       R* __restrict__ acc_a, R* __restrict__ acc_e, R* __restrict__ acc_bx,
       R* __restrict__ acc_by, R* __restrict__ acc_bz) {
    const int k = blockIdx.x * BZ + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, i = blockIdx.z;
    if (k >= g.Nz || j >= g.Ny) return;
    const int sy = g.Nz, sx = g.Ny * g.Nz;
    const int c = k + g.Nz * (j + g.Ny * i);
    const R pb = Pb[c];
    acc_a[c] += pb * Pn[c];
    acc_e[c] -= pb * div_at(Vx, Vy, Vz, c, i, j, k, g);
    const R ep = E[c] * pb;
    if (i < g.Nx - 1) {
        const R v = Vbx[c] + (E[c + sx] * Pb[c + sx] - ep);
        Vbx[c] = v;
        acc_bx[c] -= v * (Pn[c + sx] - Pn[c]);
    }
    if (j < g.Ny - 1) {
        const R v = Vby[c] + (E[c + sy] * Pb[c + sy] - ep);
        Vby[c] = v;
        acc_by[c] -= v * (Pn[c + sy] - Pn[c]);
    }
    if (k < g.Nz - 1) {
        const R v = Vbz[c] + (E[c + 1] * Pb[c + 1] - ep);
        Vbz[c] = v;
        acc_bz[c] -= v * (Pn[c + 1] - Pn[c]);
    }
}

// Pb = a Pb + D (beta Vb)
template <typename R>
__global__ void __launch_bounds__(BZ * BY)
k_apres(AdjGeom g, R* __restrict__ Pb, const R* __restrict__ Vbx, const R* __restrict__ Vby,
        const R* __restrict__ Vbz, const R* __restrict__ A, const R* __restrict__ b) {
    const int k = blockIdx.x * BZ + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, i = blockIdx.z;
    if (k >= g.Nz || j >= g.Ny) return;
    const int sy = g.Nz, sx = g.Ny * g.Nz;
    const int c = k + g.Nz * (j + g.Ny * i);
    const R h = (R)0.5 * (R)g.dtdx;
    R d = (R)0;
    if (i < g.Nx - 1) d += h * (b[c] + b[c + sx]) * Vbx[c];
    if (i > 0) d -= h * (b[c - sx] + b[c]) * Vbx[c - sx];
    if (j < g.Ny - 1) d += h * (b[c] + b[c + sy]) * Vby[c];
    if (j > 0) d -= h * (b[c - sy] + b[c]) * Vby[c - sy];
    if (k < g.Nz - 1) d += h * (b[c] + b[c + 1]) * Vbz[c];
    if (k > 0) d -= h * (b[c - 1] + b[c]) * Vbz[c - 1];
    Pb[c] = A[c] * Pb[c] + d;
}

// Chain rule from (abar, ebar, betabar) to (m, l, eta).
template <typename R>
__global__ void __launch_bounds__(BZ * BY)
k_grad(AdjGeom g, const R* acc_a, const R* acc_e, const R* acc_bx, const R* acc_by,
       const R* acc_bz, const R* b, const R* kap, const R* xi, const R* l, R* gm, R* gl, R* geta) {
    const int k = blockIdx.x * BZ + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, i = blockIdx.z;
    if (k >= g.Nz || j >= g.Ny) return;
    const int sy = g.Nz, sx = g.Ny * g.Nz;
    const int c = k + g.Nz * (j + g.Ny * i);
    const R x = xi[c], kp = kap[c];
    const R d1 = (R)1 + x;
    const R xbar = acc_a[c] * ((R)-2 / (d1 * d1)) + acc_e[c] * (-kp / (d1 * d1));
    const R kbar = acc_e[c] / d1;
    // d(kappa)/dl = kappa, d(xi)/dl = xi/2, d(xi)/dm = -xi/2
    gl[c] = kbar * kp + xbar * x * (R)0.5;
    // beta_f = (b_L + b_R)/2 dt/dx, db/dm = -b
    R fb = (R)0;
    if (i < g.Nx - 1) fb += acc_bx[c];
    if (i > 0) fb += acc_bx[c - sx];
    if (j < g.Ny - 1) fb += acc_by[c];
    if (j > 0) fb += acc_by[c - sy];
    if (k < g.Nz - 1) fb += acc_bz[c];
    if (k > 0) fb += acc_bz[c - 1];
    gm[c] = -xbar * x * (R)0.5 - fb * (R)0.5 * (R)g.dtdx * b[c];
    const bool edge = i == 0 || i == g.Nx - 1 || j == 0 || j == g.Ny - 1 || k == 0 || k == g.Nz - 1;
    // d(xi)/d(eta) = (dt/(2 dx)) sqrt(K b)
    geta[c] = edge ? xbar * (R)0.5 * (R)g.dtdx * sqrt(exp(l[c]) * b[c]) : (R)0;
}

}  // namespace

struct FdtdAdjoint {
    int Nx, Ny, Nz, N, rbytes, ckpt_every;
    double dtdx;
    hipStream_t stream;
    // coefficients
    void *b, *kap, *xi, *A, *E;
    // working forward state (ping-pong P)
    void *P[2], *V[3];
    // checkpoints (P, Vx, Vy, Vz per checkpoint)
    std::vector<void*> ck;
    int n_ck;
    // segment store: P^n (K+1) and V^{n+1/2} (K) per component
    std::vector<void*> segP, segV;
    // adjoint fields and accumulators
    void *Pb, *Vb[3], *acc[5];
    // sources / receivers
    int nsrc, slen, nrec;
    int *d_sidx, *d_ridx;
    void *d_sig, *d_rwgt;
    // receiver taps grouped by cell for the adjoint injection
    int n_icell;
    int *d_icell, *d_istart, *d_irec;
    void* d_iw;
    int n_steps;
    const void *m, *l, *eta;       // parameters of the last forward (device, caller-owned)
};

namespace {

void afree(void* p) { if (p) (void)hipFree(p); }

template <typename R>
int run_forward(FdtdAdjoint* a, const R* m, const R* l, const R* eta, int n_steps, R* traces) {
    AdjGeom g{a->Nx, a->Ny, a->Nz, a->dtdx};
    dim3 blk(BZ, BY), grd((a->Nz + BZ - 1) / BZ, (a->Ny + BY - 1) / BY, a->Nx);
    const size_t fb = (size_t)a->N * sizeof(R);
    hipLaunchKernelGGL(k_coef<R>, grd, blk, 0, a->stream, g, m, l, eta, (R*)a->b, (R*)a->kap,
                       (R*)a->xi, (R*)a->A, (R*)a->E);
    for (int q = 0; q < 2; ++q) AFDTD_CHECK(hipMemsetAsync(a->P[q], 0, fb, a->stream));
    for (int q = 0; q < 3; ++q) AFDTD_CHECK(hipMemsetAsync(a->V[q], 0, fb, a->stream));
    const int K = a->ckpt_every;
    const int nck = (n_steps + K - 1) / K;
    // (re)allocate checkpoints
    if (nck > a->n_ck) {
        for (void* p : a->ck) afree(p);
        a->ck.assign((size_t)4 * nck, nullptr);
        for (auto& p : a->ck) AFDTD_CHECK(hipMalloc(&p, fb));
        a->n_ck = nck;
    }
    int cur = 0;
    for (int n = 0; n < n_steps; ++n) {
        if (n % K == 0) {
            const int q = n / K;
            AFDTD_CHECK(hipMemcpyAsync(a->ck[4 * q], a->P[cur], fb, hipMemcpyDeviceToDevice, a->stream));
            for (int c = 0; c < 3; ++c)
                AFDTD_CHECK(hipMemcpyAsync(a->ck[4 * q + 1 + c], a->V[c], fb, hipMemcpyDeviceToDevice,
                                           a->stream));
        }
        R* P = (R*)a->P[cur];
        R* Pn = (R*)a->P[cur ^ 1];
        R* Vx = (R*)a->V[0];
        R* Vy = (R*)a->V[1];
        R* Vz = (R*)a->V[2];
        hipLaunchKernelGGL(k_fvel<R>, grd, blk, 0, a->stream, g, P, (const R*)a->b, Vx, Vy, Vz, Vx, Vy, Vz);
        hipLaunchKernelGGL(k_fpres<R>, grd, blk, 0, a->stream, g, P, Pn, Vx, Vy, Vz, (const R*)a->A,
                           (const R*)a->E);
        hipLaunchKernelGGL(k_fpost<R>, dim3(1), dim3(64), 0, a->stream, Pn, a->nsrc, a->d_sidx,
                           (const R*)a->d_sig, a->slen, n, a->nrec, a->d_ridx, (const R*)a->d_rwgt,
                           traces);
        cur ^= 1;
    }
    AFDTD_CHECK(hipGetLastError());
    AFDTD_CHECK(hipStreamSynchronize(a->stream));
    a->n_steps = n_steps;
    a->m = m; a->l = l; a->eta = eta;
    return FDTD_OK;
}

template <typename R>
int run_backward(FdtdAdjoint* a, const R* gy, R* gm, R* gl, R* geta) {
    AdjGeom g{a->Nx, a->Ny, a->Nz, a->dtdx};
    dim3 blk(BZ, BY), grd((a->Nz + BZ - 1) / BZ, (a->Ny + BY - 1) / BY, a->Nx);
    const size_t fb = (size_t)a->N * sizeof(R);
    const int K = a->ckpt_every, n_steps = a->n_steps;
    if (a->segP.empty()) {
        a->segP.assign(K + 1, nullptr);
        a->segV.assign((size_t)3 * K, nullptr);
        for (auto& p : a->segP) AFDTD_CHECK(hipMalloc(&p, fb));
        for (auto& p : a->segV) AFDTD_CHECK(hipMalloc(&p, fb));
    }
    AFDTD_CHECK(hipMemsetAsync(a->Pb, 0, fb, a->stream));
    for (int c = 0; c < 3; ++c) AFDTD_CHECK(hipMemsetAsync(a->Vb[c], 0, fb, a->stream));
    for (int c = 0; c < 5; ++c) AFDTD_CHECK(hipMemsetAsync(a->acc[c], 0, fb, a->stream));
    const int nck = (n_steps + K - 1) / K;
    for (int q = nck - 1; q >= 0; --q) {
        const int s0 = q * K, len = std::min(K, n_steps - s0);
        // recompute the segment: segP[j] = P^{s0+j}, segV[3j+c] = V^{s0+j+1/2}
        AFDTD_CHECK(hipMemcpyAsync(a->segP[0], a->ck[4 * q], fb, hipMemcpyDeviceToDevice, a->stream));
        for (int j = 0; j < len; ++j) {
            const R* Vix = j ? (const R*)a->segV[3 * (j - 1)] : (const R*)a->ck[4 * q + 1];
            const R* Viy = j ? (const R*)a->segV[3 * (j - 1) + 1] : (const R*)a->ck[4 * q + 2];
            const R* Viz = j ? (const R*)a->segV[3 * (j - 1) + 2] : (const R*)a->ck[4 * q + 3];
            R* Vox = (R*)a->segV[3 * j];
            R* Voy = (R*)a->segV[3 * j + 1];
            R* Voz = (R*)a->segV[3 * j + 2];
            hipLaunchKernelGGL(k_fvel<R>, grd, blk, 0, a->stream, g, (const R*)a->segP[j],
                               (const R*)a->b, Vix, Viy, Viz, Vox, Voy, Voz);
            hipLaunchKernelGGL(k_fpres<R>, grd, blk, 0, a->stream, g, (const R*)a->segP[j],
                               (R*)a->segP[j + 1], Vox, Voy, Voz, (const R*)a->A, (const R*)a->E);
            hipLaunchKernelGGL(k_fpost<R>, dim3(1), dim3(64), 0, a->stream, (R*)a->segP[j + 1],
                               a->nsrc, a->d_sidx, (const R*)a->d_sig, a->slen, s0 + j, a->nrec,
                               a->d_ridx, (const R*)a->d_rwgt, (R*)nullptr);
        }
        // reverse sweeps over the segment
        for (int j = len - 1; j >= 0; --j) {
            const int n = s0 + j;
            hipLaunchKernelGGL(k_inject<R>, dim3((a->n_icell + 255) / 256), dim3(256), 0, a->stream,
                               (R*)a->Pb, a->n_icell, a->d_icell, a->d_istart, a->d_irec,
                               (const R*)a->d_iw, gy + (size_t)n * a->nrec);
            hipLaunchKernelGGL(k_avel<R>, grd, blk, 0, a->stream, g, (const R*)a->Pb, (R*)a->Vb[0],
                               (R*)a->Vb[1], (R*)a->Vb[2], (const R*)a->segP[j],
                               (const R*)a->segV[3 * j], (const R*)a->segV[3 * j + 1],
                               (const R*)a->segV[3 * j + 2], (const R*)a->E, (R*)a->acc[0],
                               (R*)a->acc[1], (R*)a->acc[2], (R*)a->acc[3], (R*)a->acc[4]);
            hipLaunchKernelGGL(k_apres<R>, grd, blk, 0, a->stream, g, (R*)a->Pb, (const R*)a->Vb[0],
                               (const R*)a->Vb[1], (const R*)a->Vb[2], (const R*)a->A, (const R*)a->b);
        }
    }
    hipLaunchKernelGGL(k_grad<R>, grd, blk, 0, a->stream, g, (const R*)a->acc[0], (const R*)a->acc[1],
                       (const R*)a->acc[2], (const R*)a->acc[3], (const R*)a->acc[4], (const R*)a->b,
                       (const R*)a->kap, (const R*)a->xi, (const R*)a->l, gm, gl, geta);
    AFDTD_CHECK(hipGetLastError());
    AFDTD_CHECK(hipStreamSynchronize(a->stream));
    return FDTD_OK;
}

}  // namespace

extern "C" {

SK_API void fdtd_adjoint_destroy(FdtdAdjoint* a) {
    if (!a) return;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    afree(a->b); afree(a->kap); afree(a->xi); afree(a->A); afree(a->E);
    afree(a->P[0]); afree(a->P[1]);
    for (int c = 0; c < 3; ++c) { afree(a->V[c]); afree(a->Vb[c]); }
    afree(a->Pb);
    for (int c = 0; c < 5; ++c) afree(a->acc[c]);
    for (void* p : a->ck) afree(p);
    for (void* p : a->segP) afree(p);
    for (void* p : a->segV) afree(p);
    afree(a->d_sidx); afree(a->d_ridx); afree(a->d_sig); afree(a->d_rwgt);
    afree(a->d_icell); afree(a->d_istart); afree(a->d_irec); afree(a->d_iw);
    if (a->stream) (void)hipStreamDestroy(a->stream);
    delete a;
}

// precision_bytes: 4 (float) or 8 (double); checkpoint_every: K steps.
SK_API FdtdAdjoint* fdtd_adjoint_create(int Nx, int Ny, int Nz, double dtdx, int precision_bytes,
                                        int checkpoint_every) {
    if (Nx < 2 || Ny < 2 || Nz < 2 || dtdx <= 0.0 || checkpoint_every <= 0) return nullptr;
    if (precision_bytes != 4 && precision_bytes != 8) return nullptr;
    FdtdAdjoint* a = new FdtdAdjoint();
    a->Nx = Nx; a->Ny = Ny; a->Nz = Nz; a->N = Nx * Ny * Nz;
    a->rbytes = precision_bytes;
    a->ckpt_every = checkpoint_every;
    a->dtdx = dtdx;
    const size_t fb = (size_t)a->N * precision_bytes;
    bool ok = hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking) == hipSuccess;
    void** all[] = {&a->b, &a->kap, &a->xi, &a->A, &a->E, &a->P[0], &a->P[1], &a->V[0], &a->V[1],
                    &a->V[2], &a->Pb, &a->Vb[0], &a->Vb[1], &a->Vb[2], &a->acc[0], &a->acc[1],
                    &a->acc[2], &a->acc[3], &a->acc[4]};
    for (void** p : all) ok = ok && hipMalloc(p, fb) == hipSuccess;
    if (!ok) {
        afdtd_set_error("fdtd_adjoint_create: device allocation failed");
        fdtd_adjoint_destroy(a);
        return nullptr;
    }
    return a;
}

// Sources: cell indices and signals (REAL, n_src x n_samples, host memory).
SK_API int fdtd_adjoint_set_sources(FdtdAdjoint* a, int n_src, const int* idx, const void* sig,
                                    int n_samples) {
    if (!a || (n_src > 0 && (!idx || !sig))) return FDTD_ERR_NULL;
    for (int s = 0; s < n_src; ++s)
        if (idx[s] < 0 || idx[s] >= a->N) return FDTD_ERR_DIM;
    afree(a->d_sidx); afree(a->d_sig);
    a->d_sidx = nullptr; a->d_sig = nullptr;
    a->nsrc = n_src; a->slen = n_samples;
    if (n_src <= 0) return FDTD_OK;
    AFDTD_CHECK(hipMalloc(&a->d_sidx, sizeof(int) * n_src));
    AFDTD_CHECK(hipMalloc(&a->d_sig, (size_t)a->rbytes * n_src * n_samples));
    AFDTD_CHECK(hipMemcpy(a->d_sidx, idx, sizeof(int) * n_src, hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(a->d_sig, sig, (size_t)a->rbytes * n_src * n_samples, hipMemcpyHostToDevice));
    return FDTD_OK;
}

// Receivers: 8-tap stencils (flat cell indices, REAL weights), host memory.
SK_API int fdtd_adjoint_set_receivers(FdtdAdjoint* a, int n_rec, const int* idx8, const void* wgt8) {
    if (!a || n_rec <= 0 || !idx8 || !wgt8) return FDTD_ERR_NULL;
    for (int q = 0; q < 8 * n_rec; ++q)
        if (idx8[q] < 0 || idx8[q] >= a->N) return FDTD_ERR_DIM;
    afree(a->d_ridx); afree(a->d_rwgt);
    a->d_ridx = nullptr; a->d_rwgt = nullptr;
    AFDTD_CHECK(hipMalloc(&a->d_ridx, sizeof(int) * 8 * n_rec));
    AFDTD_CHECK(hipMalloc(&a->d_rwgt, (size_t)a->rbytes * 8 * n_rec));
    AFDTD_CHECK(hipMemcpy(a->d_ridx, idx8, sizeof(int) * 8 * n_rec, hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(a->d_rwgt, wgt8, (size_t)a->rbytes * 8 * n_rec, hipMemcpyHostToDevice));
    a->nrec = n_rec;
    // taps grouped by cell, each group in receiver-then-tap order
    const int m = 8 * n_rec;
    std::vector<int> ord(m);
    for (int q = 0; q < m; ++q) ord[q] = q;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return idx8[x] < idx8[y]; });
    std::vector<int> cell, start, rec(m);
    std::vector<char> w((size_t)m * a->rbytes);
    for (int k = 0; k < m; ++k) {
        const int q = ord[k];
        if (k == 0 || idx8[q] != idx8[ord[k - 1]]) {
            cell.push_back(idx8[q]);
            start.push_back(k);
        }
        rec[k] = q / 8;
        memcpy(&w[(size_t)k * a->rbytes], (const char*)wgt8 + (size_t)q * a->rbytes, a->rbytes);
    }
    start.push_back(m);
    afree(a->d_icell); afree(a->d_istart); afree(a->d_irec); afree(a->d_iw);
    a->d_icell = a->d_istart = a->d_irec = nullptr;
    a->d_iw = nullptr;
    a->n_icell = (int)cell.size();
    AFDTD_CHECK(hipMalloc(&a->d_icell, sizeof(int) * cell.size()));
    AFDTD_CHECK(hipMalloc(&a->d_istart, sizeof(int) * start.size()));
    AFDTD_CHECK(hipMalloc(&a->d_irec, sizeof(int) * m));
    AFDTD_CHECK(hipMalloc(&a->d_iw, w.size()));
    AFDTD_CHECK(hipMemcpy(a->d_icell, cell.data(), sizeof(int) * cell.size(), hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(a->d_istart, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(a->d_irec, rec.data(), sizeof(int) * m, hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(a->d_iw, w.data(), w.size(), hipMemcpyHostToDevice));
    return FDTD_OK;
}

// Forward with checkpoints.  m, l, eta: device arrays (N, REAL); traces:
// device array (n_steps x n_rec, REAL).  Parameters must stay alive and
// unchanged until the matching fdtd_adjoint_backward.
SK_API int fdtd_adjoint_forward(FdtdAdjoint* a, const void* m, const void* l, const void* eta,
                                int n_steps, void* traces) {
    if (!a || !m || !l || !eta || !traces || n_steps <= 0 || a->nrec <= 0) return FDTD_ERR_NULL;
    return a->rbytes == 8
               ? run_forward<double>(a, (const double*)m, (const double*)l, (const double*)eta,
                                     n_steps, (double*)traces)
               : run_forward<float>(a, (const float*)m, (const float*)l, (const float*)eta, n_steps,
                                    (float*)traces);
}

// Gradients of sum(gy * traces) with respect to m = log rho, l = log K and
// eta (device arrays, REAL).  gy: device (n_steps x n_rec).
SK_API int fdtd_adjoint_backward(FdtdAdjoint* a, const void* gy, void* gm, void* gl, void* geta) {
    if (!a || !gy || !gm || !gl || !geta || a->n_steps <= 0) return FDTD_ERR_NULL;
    return a->rbytes == 8 ? run_backward<double>(a, (const double*)gy, (double*)gm, (double*)gl,
                                                 (double*)geta)
                          : run_backward<float>(a, (const float*)gy, (float*)gm, (float*)gl,
                                                (float*)geta);
}

}  // extern "C"
