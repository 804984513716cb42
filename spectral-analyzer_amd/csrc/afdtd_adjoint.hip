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
//
// Each step is one sweep in either direction: a thread owns cell c and the
// three faces on its + side, and also evaluates the three faces on its -
// side (owned by the neighbours) with the owners' expressions, so it has all
// six faces of its divergence without a second launch.  The recomputed faces
// are bit-identical to the owners' (same operands in the same order), and
// the old/new fields are separate buffers (no thread reads a value another
// thread of the same sweep writes).  The per-face accumulators betabar are
// folded into one per-cell sum over the cell's six faces, which is what the
// chain rule for log rho needs; abar is only accumulated on domain-edge cells,
// the only cells whose gradient it reaches (xi = 0 elsewhere).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "afdtd_internal.h"

namespace {

constexpr int BZ = 64, BY = 4;

// Cell of this thread in the one-dimensional launches of the two step
// sweeps.  Consecutive workgroups go to the eight XCDs in turn; the remap
// gives each XCD a contiguous range of (z-block, y-block, x-plane) tiles in
// plane order, so the +-x neighbours a tile reads were fetched into the same
// XCD's L2 by the tiles of the planes next to it.
struct Cell3 {
    int i, j, k;
};
__device__ __forceinline__ Cell3 sweep_cell(int Ny, int Nz) {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int x = b & 7, per = nb >> 3, rem = nb & 7;
    const int L = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + (b >> 3);
    const int gz = (Nz + BZ - 1) / BZ, gy = (Ny + BY - 1) / BY;
    const int bz = L % gz, t = L / gz;
    return Cell3{t / gy, (t % gy) * BY + (int)threadIdx.y, bz * BZ + (int)threadIdx.x};
}
constexpr int FUSED_SRC_MAX = 16;   // sources added inside the forward sweep

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

// One forward step: V_out = V_in - beta (G P_in) on the cell's + faces
// (written) and - faces (the owners' values, recomputed), then
// P_out = a P_in - e (D V_out).  Faces that do not exist are written as 0.
template <typename R>
__global__ void __launch_bounds__(BZ * BY)
k_fstep(AdjGeom g, const R* __restrict__ P, const R* __restrict__ b, const R* __restrict__ Vxi,
        const R* __restrict__ Vyi, const R* __restrict__ Vzi, R* __restrict__ Vxo,
        R* __restrict__ Vyo, R* __restrict__ Vzo, R* __restrict__ Pout, const R* __restrict__ A,
        const R* __restrict__ E, int nsrc, const int* __restrict__ sidx,
        const R* __restrict__ sig, int slen, int n) {
    const Cell3 cl = sweep_cell(g.Ny, g.Nz);
    const int i = cl.i, j = cl.j, k = cl.k;
    if (k >= g.Nz || j >= g.Ny) return;
    const int sy = g.Nz, sx = g.Ny * g.Nz;
    const int c = k + g.Nz * (j + g.Ny * i);
    const R h = (R)0.5 * (R)g.dtdx;
    const R pc = P[c], bc = b[c];
    // face between cells L and L + s: V_in - h (b_L + b_{L+s}) (P_{L+s} - P_L)
    auto face = [&](const R* Vi, int L, int s) __attribute__((always_inline)) {
        return Vi[L] - h * (b[L] + b[L + s]) * (P[L + s] - P[L]);
    };
    const bool xp = i < g.Nx - 1, yp = j < g.Ny - 1, zp = k < g.Nz - 1;
    const R vxp = xp ? Vxi[c] - h * (bc + b[c + sx]) * (P[c + sx] - pc) : (R)0;
    const R vyp = yp ? Vyi[c] - h * (bc + b[c + sy]) * (P[c + sy] - pc) : (R)0;
    const R vzp = zp ? Vzi[c] - h * (bc + b[c + 1]) * (P[c + 1] - pc) : (R)0;
    Vxo[c] = vxp;
    Vyo[c] = vyp;
    Vzo[c] = vzp;
    R d = (R)0;
    if (xp) d += vxp;
    if (i > 0) d -= face(Vxi, c - sx, sx);
    if (yp) d += vyp;
    if (j > 0) d -= face(Vyi, c - sy, sy);
    if (zp) d += vzp;
    if (k > 0) d -= face(Vzi, c - 1, 1);
    R out = A[c] * pc - E[c] * d;
    // point sources of this step, added in source order (nsrc = 0: added
    // by k_fpost instead)
    if (n < slen)
        for (int q = 0; q < nsrc; ++q)
            if (sidx[q] == c) out += sig[(size_t)q * slen + n];
    Pout[c] = out;
}

// Sources after the pressure update (when k_fstep did not add them);
// receivers record y_n (forward only).
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

// One reverse step (after the receiver injection into Pb_in):
//   Vb_out = Vb_in + G (e Pb_in) on the cell's + faces (written) and - faces
//   (recomputed), Pb_out = a Pb_in + D (beta Vb_out), and the accumulations
//   abar += Pb P^n (edge cells), ebar -= Pb (D V^{n+1/2}),
//   fbar -= sum over the cell's faces of Vb_out (G P^n),
// then adds the receiver injection R^T gy of the next reverse step into
// Pb_out (gy null: none; cmap[c] = the cell's entry in the tap-cell lists or
// -1).
template <typename R>
__global__ void __launch_bounds__(BZ * BY)
k_astep(AdjGeom g, const R* __restrict__ Pb, R* __restrict__ Pbo, const R* __restrict__ Vbxi,
        const R* __restrict__ Vbyi, const R* __restrict__ Vbzi, R* __restrict__ Vbxo,
        R* __restrict__ Vbyo, R* __restrict__ Vbzo, const R* __restrict__ Pn,
        const R* __restrict__ Vx, const R* __restrict__ Vy, const R* __restrict__ Vz,
        const R* __restrict__ A, const R* __restrict__ E, const R* __restrict__ b,
        R* __restrict__ acc_a, R* __restrict__ acc_e, R* __restrict__ acc_f,
        const int* __restrict__ cmap, const int* __restrict__ start, const int* __restrict__ rec,
        const R* __restrict__ w, const R* __restrict__ gy) {
    const Cell3 cl = sweep_cell(g.Ny, g.Nz);
    const int i = cl.i, j = cl.j, k = cl.k;
    if (k >= g.Nz || j >= g.Ny) return;
    const int sy = g.Nz, sx = g.Ny * g.Nz;
    const int c = k + g.Nz * (j + g.Ny * i);
    const R h = (R)0.5 * (R)g.dtdx;
    const R pb = Pb[c], pnc = Pn[c], bc = b[c];
    const bool edge = i == 0 || i == g.Nx - 1 || j == 0 || j == g.Ny - 1 || k == 0 || k == g.Nz - 1;
    if (edge) acc_a[c] += pb * pnc;
    acc_e[c] -= pb * div_at(Vx, Vy, Vz, c, i, j, k, g);
    const R ep = E[c] * pb;
    R f = acc_f[c], d = (R)0;
    if (i < g.Nx - 1) {
        const R v = Vbxi[c] + (E[c + sx] * Pb[c + sx] - ep);
        Vbxo[c] = v;
        f -= v * (Pn[c + sx] - pnc);
        d += h * (bc + b[c + sx]) * v;
    }
    if (i > 0) {
        const R v = Vbxi[c - sx] + (ep - E[c - sx] * Pb[c - sx]);
        f -= v * (pnc - Pn[c - sx]);
        d -= h * (b[c - sx] + bc) * v;
    }
    if (j < g.Ny - 1) {
        const R v = Vbyi[c] + (E[c + sy] * Pb[c + sy] - ep);
        Vbyo[c] = v;
        f -= v * (Pn[c + sy] - pnc);
        d += h * (bc + b[c + sy]) * v;
    }
    if (j > 0) {
        const R v = Vbyi[c - sy] + (ep - E[c - sy] * Pb[c - sy]);
        f -= v * (pnc - Pn[c - sy]);
        d -= h * (b[c - sy] + bc) * v;
    }
    if (k < g.Nz - 1) {
        const R v = Vbzi[c] + (E[c + 1] * Pb[c + 1] - ep);
        Vbzo[c] = v;
        f -= v * (Pn[c + 1] - pnc);
        d += h * (bc + b[c + 1]) * v;
    }
    if (k > 0) {
        const R v = Vbzi[c - 1] + (ep - E[c - 1] * Pb[c - 1]);
        f -= v * (pnc - Pn[c - 1]);
        d -= h * (b[c - 1] + bc) * v;
    }
    acc_f[c] = f;
    R out = A[c] * pb + d;
    // the next reverse step's receiver injection (gy of the step before; as
    // k_inject: this cell's taps in receiver-then-tap order)
    if (gy) {
        const int u = cmap[c];
        if (u >= 0)
            for (int q = start[u]; q < start[u + 1]; ++q) out += w[q] * gy[rec[q]];
    }
    Pbo[c] = out;
}

// Chain rule from (abar, ebar, betabar) to (m, l, eta).
template <typename R>
__global__ void __launch_bounds__(BZ * BY)
k_grad(AdjGeom g, const R* acc_a, const R* acc_e, const R* acc_f, const R* b, const R* kap,
       const R* xi, const R* l, R* gm, R* gl, R* geta) {
    const int k = blockIdx.x * BZ + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, i = blockIdx.z;
    if (k >= g.Nz || j >= g.Ny) return;
    const int c = k + g.Nz * (j + g.Ny * i);
    const R x = xi[c], kp = kap[c];
    const R d1 = (R)1 + x;
    const R xbar = acc_a[c] * ((R)-2 / (d1 * d1)) + acc_e[c] * (-kp / (d1 * d1));
    const R kbar = acc_e[c] / d1;
    // d(kappa)/dl = kappa, d(xi)/dl = xi/2, d(xi)/dm = -xi/2
    gl[c] = kbar * kp + xbar * x * (R)0.5;
    // beta_f = (b_L + b_R)/2 dt/dx, db/dm = -b; fb: betabar summed over the
    // cell's faces
    const R fb = acc_f[c];
    gm[c] = -xbar * x * (R)0.5 - fb * (R)0.5 * (R)g.dtdx * b[c];
    const bool edge = i == 0 || i == g.Nx - 1 || j == 0 || j == g.Ny - 1 || k == 0 || k == g.Nz - 1;
    // d(xi)/d(eta) = (dt/(2 dx)) sqrt(K b)
    geta[c] = edge ? xbar * (R)0.5 * (R)g.dtdx * sqrt(exp(l[c]) * b[c]) : (R)0;
}

}  // namespace

struct FdtdAdjoint {
    int Nx, Ny, Nz, N, rbytes, ckpt_every;
    double dtdx;
    hipStream_t stream;     // forward, reverse sweeps
    hipStream_t stream2;    // segment recompute, overlapped with the reverse sweeps
    hipEvent_t ev_seg[2];   // segment store s recomputed
    hipEvent_t ev_rev[2];   // reverse sweeps done reading segment store s
    // coefficients
    void *b, *kap, *xi, *A, *E;
    // working forward state (ping-pong P and V)
    void *P[2], *V[2][3];
    // checkpoints (P, Vx, Vy, Vz per checkpoint)
    std::vector<void*> ck;
    int n_ck;
    // two segment stores (alternating by segment): P^n (K+1) and V^{n+1/2}
    // (K) per component
    std::vector<void*> segP[2], segV[2];
    // adjoint fields (ping-pong) and accumulators abar, ebar, fbar
    void *Pb[2], *Vb[2][3], *acc[3];
    // sources / receivers
    int nsrc, slen, nrec;
    int *d_sidx, *d_ridx;
    void *d_sig, *d_rwgt;
    // receiver taps grouped by cell for the adjoint injection
    int n_icell;
    int *d_icell, *d_istart, *d_irec, *d_icmap;
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
    const dim3 grd1(grd.x * grd.y * grd.z);
    const size_t fb = (size_t)a->N * sizeof(R);
    hipLaunchKernelGGL(k_coef<R>, grd, blk, 0, a->stream, g, m, l, eta, (R*)a->b, (R*)a->kap,
                       (R*)a->xi, (R*)a->A, (R*)a->E);
    for (int q = 0; q < 2; ++q) AFDTD_CHECK(hipMemsetAsync(a->P[q], 0, fb, a->stream));
    for (int q = 0; q < 3; ++q) AFDTD_CHECK(hipMemsetAsync(a->V[0][q], 0, fb, a->stream));
    const int K = a->ckpt_every;
    const int nck = (n_steps + K - 1) / K;
    // (re)allocate checkpoints
    if (nck > a->n_ck) {
        for (void* p : a->ck) afree(p);
        a->ck.assign((size_t)4 * nck, nullptr);
        for (auto& p : a->ck) AFDTD_CHECK(hipMalloc(&p, fb));
        a->n_ck = nck;
    }
    // few sources are added inside the sweep, many by k_fpost
    const int fsrc = a->nsrc <= FUSED_SRC_MAX ? a->nsrc : 0, psrc = a->nsrc - fsrc;
    int cur = 0;
    for (int n = 0; n < n_steps; ++n) {
        void* const* V = a->V[cur];
        void* const* Vn = a->V[cur ^ 1];
        if (n % K == 0) {
            const int q = n / K;
            AFDTD_CHECK(hipMemcpyAsync(a->ck[4 * q], a->P[cur], fb, hipMemcpyDeviceToDevice, a->stream));
            for (int c = 0; c < 3; ++c)
                AFDTD_CHECK(hipMemcpyAsync(a->ck[4 * q + 1 + c], V[c], fb, hipMemcpyDeviceToDevice,
                                           a->stream));
        }
        R* Pn = (R*)a->P[cur ^ 1];
        hipLaunchKernelGGL(k_fstep<R>, grd1, blk, 0, a->stream, g, (const R*)a->P[cur], (const R*)a->b,
                           (const R*)V[0], (const R*)V[1], (const R*)V[2], (R*)Vn[0], (R*)Vn[1],
                           (R*)Vn[2], Pn, (const R*)a->A, (const R*)a->E, fsrc, a->d_sidx,
                           (const R*)a->d_sig, a->slen, n);
        hipLaunchKernelGGL(k_fpost<R>, dim3(1), dim3(64), 0, a->stream, Pn, psrc, a->d_sidx,
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
    const dim3 grd1(grd.x * grd.y * grd.z);
    const size_t fb = (size_t)a->N * sizeof(R);
    const int K = a->ckpt_every, n_steps = a->n_steps;
    if (a->segP[0].empty()) {
        for (int u = 0; u < 2; ++u) {
            a->segP[u].assign(K + 1, nullptr);
            a->segV[u].assign((size_t)3 * K, nullptr);
            for (auto& p : a->segP[u]) AFDTD_CHECK(hipMalloc(&p, fb));
            for (auto& p : a->segV[u]) AFDTD_CHECK(hipMalloc(&p, fb));
        }
    }
    // both Vb buffers start at 0: faces that do not exist are never written
    AFDTD_CHECK(hipMemsetAsync(a->Pb[0], 0, fb, a->stream));
    for (int u = 0; u < 2; ++u)
        for (int c = 0; c < 3; ++c) AFDTD_CHECK(hipMemsetAsync(a->Vb[u][c], 0, fb, a->stream));
    for (int c = 0; c < 3; ++c) AFDTD_CHECK(hipMemsetAsync(a->acc[c], 0, fb, a->stream));
    const int nck = (n_steps + K - 1) / K;
    const int fsrc = a->nsrc <= FUSED_SRC_MAX ? a->nsrc : 0;
    // Segment q is recomputed into store q % 2 on stream2 while the reverse
    // sweeps of segment q + 1 run on stream over the other store.
    auto recompute = [&](int q) -> int {
        const int u = q & 1, s0 = q * K, len = std::min(K, n_steps - s0);
        std::vector<void*>& sP = a->segP[u];
        std::vector<void*>& sV = a->segV[u];
        if (q + 2 < nck) AFDTD_CHECK(hipStreamWaitEvent(a->stream2, a->ev_rev[u], 0));
        // segP[j] = P^{s0+j}, segV[3j+c] = V^{s0+j+1/2}
        AFDTD_CHECK(hipMemcpyAsync(sP[0], a->ck[4 * q], fb, hipMemcpyDeviceToDevice, a->stream2));
        for (int j = 0; j < len; ++j) {
            const R* Vix = j ? (const R*)sV[3 * (j - 1)] : (const R*)a->ck[4 * q + 1];
            const R* Viy = j ? (const R*)sV[3 * (j - 1) + 1] : (const R*)a->ck[4 * q + 2];
            const R* Viz = j ? (const R*)sV[3 * (j - 1) + 2] : (const R*)a->ck[4 * q + 3];
            hipLaunchKernelGGL(k_fstep<R>, grd1, blk, 0, a->stream2, g, (const R*)sP[j],
                               (const R*)a->b, Vix, Viy, Viz, (R*)sV[3 * j], (R*)sV[3 * j + 1],
                               (R*)sV[3 * j + 2], (R*)sP[j + 1], (const R*)a->A, (const R*)a->E,
                               fsrc, a->d_sidx, (const R*)a->d_sig, a->slen, s0 + j);
            if (a->nsrc > fsrc)
                hipLaunchKernelGGL(k_fpost<R>, dim3(1), dim3(64), 0, a->stream2, (R*)sP[j + 1],
                                   a->nsrc, a->d_sidx, (const R*)a->d_sig, a->slen, s0 + j, a->nrec,
                                   a->d_ridx, (const R*)a->d_rwgt, (R*)nullptr);
        }
        AFDTD_CHECK(hipEventRecord(a->ev_seg[u], a->stream2));
        return FDTD_OK;
    };
    int cur = 0;
    // receiver injection of the last step; each reverse sweep adds the next one
    hipLaunchKernelGGL(k_inject<R>, dim3((a->n_icell + 255) / 256), dim3(256), 0, a->stream,
                       (R*)a->Pb[0], a->n_icell, a->d_icell, a->d_istart, a->d_irec,
                       (const R*)a->d_iw, gy + (size_t)(n_steps - 1) * a->nrec);
    int rc = recompute(nck - 1);
    if (rc != FDTD_OK) return rc;
    for (int q = nck - 1; q >= 0; --q) {
        const int u = q & 1, s0 = q * K, len = std::min(K, n_steps - s0);
        if (q > 0 && (rc = recompute(q - 1)) != FDTD_OK) return rc;
        AFDTD_CHECK(hipStreamWaitEvent(a->stream, a->ev_seg[u], 0));
        std::vector<void*>& sP = a->segP[u];
        std::vector<void*>& sV = a->segV[u];
        for (int j = len - 1; j >= 0; --j) {
            const int n = s0 + j;
            const R* Pb = (const R*)a->Pb[cur];
            void* const* Vb = a->Vb[cur];
            void* const* Vbn = a->Vb[cur ^ 1];
            const R* gyn = n > 0 ? gy + (size_t)(n - 1) * a->nrec : nullptr;
            hipLaunchKernelGGL(k_astep<R>, grd1, blk, 0, a->stream, g, Pb, (R*)a->Pb[cur ^ 1],
                               (const R*)Vb[0], (const R*)Vb[1], (const R*)Vb[2], (R*)Vbn[0],
                               (R*)Vbn[1], (R*)Vbn[2], (const R*)sP[j], (const R*)sV[3 * j],
                               (const R*)sV[3 * j + 1], (const R*)sV[3 * j + 2], (const R*)a->A,
                               (const R*)a->E, (const R*)a->b, (R*)a->acc[0], (R*)a->acc[1],
                               (R*)a->acc[2], a->d_icmap, a->d_istart, a->d_irec,
                               (const R*)a->d_iw, gyn);
            cur ^= 1;
        }
        AFDTD_CHECK(hipEventRecord(a->ev_rev[u], a->stream));
    }
    hipLaunchKernelGGL(k_grad<R>, grd, blk, 0, a->stream, g, (const R*)a->acc[0], (const R*)a->acc[1],
                       (const R*)a->acc[2], (const R*)a->b, (const R*)a->kap, (const R*)a->xi,
                       (const R*)a->l, gm, gl, geta);
    AFDTD_CHECK(hipGetLastError());
    AFDTD_CHECK(hipStreamSynchronize(a->stream2));
    AFDTD_CHECK(hipStreamSynchronize(a->stream));
    return FDTD_OK;
}

}  // namespace

extern "C" {

SK_API void fdtd_adjoint_destroy(FdtdAdjoint* a) {
    if (!a) return;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    afree(a->b); afree(a->kap); afree(a->xi); afree(a->A); afree(a->E);
    for (int u = 0; u < 2; ++u) {
        afree(a->P[u]);
        afree(a->Pb[u]);
        for (int c = 0; c < 3; ++c) { afree(a->V[u][c]); afree(a->Vb[u][c]); }
    }
    for (int c = 0; c < 3; ++c) afree(a->acc[c]);
    for (void* p : a->ck) afree(p);
    for (int u = 0; u < 2; ++u) {
        for (void* p : a->segP[u]) afree(p);
        for (void* p : a->segV[u]) afree(p);
    }
    afree(a->d_sidx); afree(a->d_ridx); afree(a->d_sig); afree(a->d_rwgt);
    afree(a->d_icell); afree(a->d_istart); afree(a->d_irec); afree(a->d_iw); afree(a->d_icmap);
    if (a->stream2) (void)hipStreamSynchronize(a->stream2);
    for (int u = 0; u < 2; ++u) {
        if (a->ev_seg[u]) (void)hipEventDestroy(a->ev_seg[u]);
        if (a->ev_rev[u]) (void)hipEventDestroy(a->ev_rev[u]);
    }
    if (a->stream2) (void)hipStreamDestroy(a->stream2);
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
    ok = ok && hipStreamCreateWithFlags(&a->stream2, hipStreamNonBlocking) == hipSuccess;
    for (int u = 0; u < 2; ++u) {
        ok = ok && hipEventCreateWithFlags(&a->ev_seg[u], hipEventDisableTiming) == hipSuccess;
        ok = ok && hipEventCreateWithFlags(&a->ev_rev[u], hipEventDisableTiming) == hipSuccess;
    }
    void** all[] = {&a->b, &a->kap, &a->xi, &a->A, &a->E, &a->P[0], &a->P[1], &a->V[0][0],
                    &a->V[0][1], &a->V[0][2], &a->V[1][0], &a->V[1][1], &a->V[1][2], &a->Pb[0],
                    &a->Pb[1], &a->Vb[0][0], &a->Vb[0][1], &a->Vb[0][2], &a->Vb[1][0],
                    &a->Vb[1][1], &a->Vb[1][2], &a->acc[0], &a->acc[1], &a->acc[2]};
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
    afree(a->d_icell); afree(a->d_istart); afree(a->d_irec); afree(a->d_iw); afree(a->d_icmap);
    a->d_icell = a->d_istart = a->d_irec = a->d_icmap = nullptr;
    a->d_iw = nullptr;
    a->n_icell = (int)cell.size();
    std::vector<int> cmap((size_t)a->N, -1);
    for (int u = 0; u < a->n_icell; ++u) cmap[cell[u]] = u;
    AFDTD_CHECK(hipMalloc(&a->d_icmap, sizeof(int) * cmap.size()));
    AFDTD_CHECK(hipMemcpy(a->d_icmap, cmap.data(), sizeof(int) * cmap.size(), hipMemcpyHostToDevice));
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
