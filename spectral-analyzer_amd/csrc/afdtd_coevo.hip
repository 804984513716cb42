// afdtd_coevo.hip — the instrument co-evolver on the device: strings, shared
// neck mode, bridge impedances, plate loads, saddle feedback, modal tracker,
// pickups and microphones, coupled to the uniform body FDTD of this library
// (reference: csrc/kernels/acoustic_coevolver.cpp, coevolver_create :987,
// _step_one_sample :1910-2226, coevolver_step :2229).
//
// Per body sub-step the host enqueues ONE single-workgroup kernel, k_coevo,
// and then the body FDTD step (plate step, fused sweep).  k_coevo first
// finishes the previous sub-step (saddle gather from the plate, modal
// projection every modal_stride sub-steps, and at a sample's end the
// pickups, microphones and output rings), then runs this sub-step's string
// side: neck mode, every string's own sub-steps (tension from the geometric
// stretch, both polarisations, shape checks), bridge impedances and the
// bridge / neck loads into the plate's force accumulator.  A block of
// samples is one enqueue loop with one synchronisation at its end; the
// strings live in the workgroup's LDS during a kernel (one wavefront per
// string) and in HBM between kernels.
//
// Arithmetic follows the reference expression by expression (the library is
// built with -ffp-contract=off and IEEE division / square root).  Sums the
// reference forms with plain loops keep their order; sums it forms with
// Eigen reductions (tension arc length, modal and pickup GEMVs, the
// bridge-drive GEMV) are taken term by term in index order, one lane each.
// Set-up (frames, B-kernels, Gaussian kernels, modal basis, pluck ramps) is
// host code with the C library's float functions.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/afdtd_coevolver.h"
#include "../../include/afdtd_pressure_backend.hpp"
#include "afdtd_internal.h"

namespace {

// stage label of the last coevolver_create_amr failure (acoustic_coevolver.cpp:31-34)
thread_local std::string g_amr_err;

constexpr int kModes = 32;          // acoustic_coevolver.cpp:40
constexpr int kMaxSub = 16;         // :41
constexpr int kRing = 4096;         // :42
constexpr float kMaxSlope = 0.25f;  // :43
constexpr float kMaxStretch = 0.0015f;
constexpr float kPi = 3.14159265358979323846f;
constexpr int kMaxStrings = 64;
constexpr int kMaxNodes = 4096;
constexpr int kThreads = 1024;
constexpr int kChunk = 1024;        // samples per enqueue loop
constexpr int kLdsLimit = 160 * 1024;

// Per-string constants on the device.
struct CoStr {
    int n_nodes, node_off, seg_off, n_sub;
    float ds, L, T0, EA, mu, dt_s, biharm, gamma_eff;
    float b_m, b_k, b_c, inv_dt_s;
};

// Per-string state words (device): T_eff, saddle_w, saddle_w_prev, bridge_w, bridge_w_prev,
// and the arc length of the current shape with its valid flag (the last
// shape check's sum, which is the next tension's sum: same shape, same terms).
enum { kTeff = 0, kSad = 1, kSadP = 2, kBr = 3, kBrP = 4, kArc = 5, kArcOk = 6, kSw = 8 };

struct CoConst {
    int ns, npk, nmic, nmag, tot_seg, tot_nodes, nb, nbgrp, nn, nngrp, neck_on, any_vel;
    int Nplate, clear_ext;
    int amr;                                // body: 0 uniform FDTD, 1 AMR stepper
    int defer;                              // output sums run on idle wavefronts (see k_coevo)
    const float* npre;                      // neck plate sum formed by the body sweep: sum, tag
    int sad_early;                          // saddle gather in the run phase (n_strings <= 12)
    int sbk_lds;                            // per-string bridge kernels staged in LDS
    int anp, anv;                           // AMR mic stencil entries: pressure, velocity
    int st_ok;                              // status[0] when nothing failed (0 / INT_MAX)
    int lds_str;                            // LDS words before the neck products
    float dt_b, inv_dx2, rhoc;
    float n_m, n_c, n_k, nb_k, nb_c;      // neck mass, total damping / stiffness, joint terms
    const CoStr* str;
    float* sw;                              // [ns][kSw] state words
    float* u;                               // [4][tot_nodes]: u0, u1, up0, up1
    float* neck;                            // neck_w, neck_w_prev, neck_plate_w_prev
    const int* nidx;                        // neck sample list
    const float* nwgt;
    const int* ngrp;                        // neck force groups (start | node | entry)
    const float* sbk;                       // [ns][nb] per-string bridge kernels
    const int* bgrp;                        // bridge force groups of the body FDTD
    const int* sad_idx;
    const float* sad_wgt;
    const float* basis;                     // per string [node][mode] at node_off * kModes
    float* modal;                           // [ns][2 * kModes]
    const float* K0;                        // [nmag][tot_seg]
    const float* K1;
    const int* magmap;
    const int* pk_type;
    const unsigned* pk_mask;
    const float* pk_sens;
    const int* mPi;                         // [nmic][8] pressure taps
    const float* mPw;
    const int* mVi;                         // [3][nmic][8] velocity taps (padded layout)
    const float* mVw;
    const float* mic_par;                   // [nmic][6]: polar_a, polar_b, gain, axis
    float* pk_out;                          // [npk][kRing]
    float* mic_out;                         // [nmic][kRing]
    const float* blk_sub;                   // [samples][ns] per-sub-step force
    const int* blk_node;                    // [samples][ns] its node
    const float* blk_dval;                  // [samples][ns] direct drive (block-drive form)
    const int* blk_dnode;                   // [ns]
    int* status;                            // body status words
    // AMR body: bridge plate weights and the mic stencils (CSR over mics)
    const float* bwgt;
    const int* amp_start;
    const int* amp_cell;
    const float* amp_w;
    const int* amv_start;
    const int* amv_face;
    const float* amv_w[3];
};

struct CoLaunch {
    int sample, sub, step;
    int npre;                               // the previous body step formed the neck plate sum
    int post, modal, sample_end, out_pos, run;
    const float* w;                         // current plate displacement
    float* ext;                             // plate force accumulator
    const float* P;
    const float* V[3];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

__device__ __forceinline__ bool wave_any(bool b) { return __ballot(b) != 0ull; }

// A failure found by this kernel before body step a.step: the uniform
// body's status becomes {FDTD_STATUS_HOOK, step}; the AMR body's hook word
// (status[3]) takes the step and its failing-step word drops below it, so
// that the step and everything after it is skipped.
__device__ __forceinline__ void fail_here(const CoConst& c, const CoLaunch& a, int* sflag) {
    *sflag = 1;
    if (c.amr) {
        if (atomicCAS(&c.status[3], INT_MAX, a.step) == INT_MAX) atomicMin(&c.status[0], a.step - 1);
    } else if (atomicCAS(&c.status[0], 0, FDTD_STATUS_HOOK) == 0) {
        c.status[1] = a.step;
    }
}

// acc + x[0] + x[1] + ... + x[n-1], one lane, in index order.  The terms
// come from LDS in blocks of sixteen, two blocks in flight: the next block
// is loaded while the adds of the current one run, so the dependent adds
// are not held up by the load latency.
__device__ __forceinline__ void load16(float* r, const float* x) {
#pragma unroll
    for (int q = 0; q < 16; ++q) r[q] = x[q];
}
__device__ __forceinline__ float add16(float acc, const float* r) {
#pragma unroll
    for (int q = 0; q < 16; ++q) acc += r[q];
    return acc;
}
__device__ __forceinline__ float seq_sum(const float* x, int n, float acc) {
    const int nb = n >> 4;                      // full blocks of sixteen
    float a[16], b[16];
    if (nb > 0) {
        load16(a, x);
        int k = 1;
        // (the scheduling barriers keep each block's loads ahead of the
        // previous block's adds; the compiler would otherwise sink them)
        for (; k + 1 < nb; k += 2) {            // blocks k (into b) and k + 1 (into a)
            load16(b, x + 16 * k);
            __builtin_amdgcn_sched_barrier(0);
            acc = add16(acc, a);
            load16(a, x + 16 * (k + 1));
            __builtin_amdgcn_sched_barrier(0);
            acc = add16(acc, b);
        }
        if (k < nb) {                           // block k is the last one
            load16(b, x + 16 * k);
            __builtin_amdgcn_sched_barrier(0);
            acc = add16(acc, a);
            acc = add16(acc, b);
        } else {
            acc = add16(acc, a);
        }
    }
    for (int i = nb * 16; i < n; ++i) acc += x[i];
    return acc;
}

// One polarisation step of a string, interior nodes by lane
// (acoustic_coevolver.cpp:595-663): returns true when a value is not finite.
__device__ bool string_pol(const float* __restrict__ u, const float* __restrict__ up,
                           float* __restrict__ ut, const float* fe, int N, float c2, float omh,
                           float inv, float biharm, float endR, float endL, int lane) {
    bool bad = false;
    for (int i = lane; i < N; i += 64) {
        float x;
        if (i == N - 1) {
            x = endR;
        } else if (i == 0) {
            x = endL;
        } else {
            const float f = fe ? fe[i] : 0.0f;
            x = 2.0f * u[i] - omh * up[i] + c2 * (u[i + 1] - 2.0f * u[i] + u[i - 1]) + f;
            if (N > 4 && biharm > 0.0f && i >= 2 && i <= N - 3)
                x -= biharm * (u[i - 2] - 4.0f * u[i - 1] + 6.0f * u[i] - 4.0f * u[i + 1] + u[i + 2]);
            x *= inv;
            if (!__builtin_isfinite(x)) bad = true;
        }
        ut[i] = x;
    }
    return wave_any(bad);
}

// Tension from the arc length (acoustic_coevolver.cpp:576-592): strain,
// T0 + EA strain and the CFL check.  One lane; returns 1 on failure.
__device__ __forceinline__ int tension_arc(const CoStr& s, float arc, float* teff_out) {
    int tbad = 0;
    float strain = (arc - s.L) / s.L;
    if (!__builtin_isfinite(strain)) tbad = 1;
    if (strain < 0.0f) strain = 0.0f;
    const float teff = s.T0 + s.EA * strain;
    if (!__builtin_isfinite(teff) || teff <= 0.0f) tbad = 1;
    const float cfl = sqrtf(teff / s.mu) * s.dt_s / s.ds;
    if (!__builtin_isfinite(cfl) || cfl * cfl > 0.9025f) tbad = 1;
    *teff_out = teff;
    return tbad;
}

// Geometric tension of a string from its current displacements (u0, u1):
// arc length in index order, then tension_arc (acoustic_coevolver.cpp:
// 563-593).  Lane 0 stores T_eff; returns true (wave-uniform) on failure.
__device__ bool tension_of(const float* u0, const float* u1, float* scr, int N, const CoStr& s,
                           int lane, float* teff_out) {
    const float ds2 = s.ds * s.ds;
    bool bad = false;
    for (int i = lane; i < N - 1; i += 64) {
        const float d0 = u0[i + 1] - u0[i];
        const float d1 = u1[i + 1] - u1[i];
        if (!__builtin_isfinite(d0) || !__builtin_isfinite(d1)) bad = true;
        scr[i] = sqrtf(ds2 + d0 * d0 + d1 * d1);
    }
    wave_sync();
    if (wave_any(bad)) return true;
    int tbad = 0;
    if (lane == 0) tbad = tension_arc(s, seq_sum(scr, N - 1, 0.0f), teff_out);
    return __shfl(tbad, 0) != 0;
}

// LDS: 8 words per string node (six polarisation buffers, the pol-0 force,
// a sum scratch), then neck products [nn], fpl[64], drive[nb], flags[8],
// per-string T_eff and u0[1] [2 * ns], the current sub-step's T_eff [ns].  The post phase reuses the string
// words as scratch (modal staging 2T, pickup velocities and products 4S).
__global__ __launch_bounds__(kThreads) void k_coevo(CoConst c, CoLaunch a) {
    if (c.status[0] != c.st_ok) return;
    extern __shared__ float lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float* nprod = lds + c.lds_str;
    float* fpl = nprod + c.nn;
    float* drive = fpl + kMaxStrings;
    int* flag = (int*)(drive + c.nb);          // [0] failure, [1..2] neck_w / neck force, [3..4] sums
    float* fsT = (float*)(flag + 8);
    float* fsU = fsT + c.ns;
    float* teff0 = fsU + c.ns;                 // T_eff of each string's current sub-step
    float* swl = teff0 + c.ns;                 // [ns][kSw] state words for this launch
    float* psub = swl + c.ns * kSw;            // [ns] this sample's per-sub-step force
    int* pnode = (int*)(psub + c.ns);          // [ns] its node
    float* sbkl = (float*)(pnode + c.ns);      // [ns][nb] bridge kernels (c.sbk_lds)
    float* dprod = sbkl + (c.sbk_lds ? c.ns * c.nb : 0);   // deferred output products (c.defer)
    if (tid == 0) flag[0] = 0;
    // At a sample's end the products of the pickup GEMV rows (and of the AMR
    // mic stencils) are formed here, but their one-lane sums run in the run
    // phase on two wavefronts that carry no string, beside the neck mode and
    // the strings' loads, instead of holding up every wavefront here.
    const bool dfr = c.defer && a.run;

    // ── finish the previous sub-step (:2105-2125, and :2129-2223 at a sample's end) ──
    if (a.post) {
        if (tid < c.ns && !(c.sad_early && a.run)) {
            const int* id = c.sad_idx + 4 * tid;
            const float* g = c.sad_wgt + 4 * tid;
            // the AMR sampler's sum starts from +0 (acoustic_amr.cpp:1600-1616)
            const float t0 = c.amr ? 0.0f + g[0] * a.w[id[0]] : g[0] * a.w[id[0]];
            c.sw[tid * kSw + kSad] = t0 + g[1] * a.w[id[1]] + g[2] * a.w[id[2]] + g[3] * a.w[id[3]];
        }
        if (a.modal) {
            for (int i = tid; i < c.tot_nodes; i += kThreads) {
                lds[i] = c.u[i];
                lds[c.tot_nodes + i] = c.u[2 * c.tot_nodes + i];
            }
            __syncthreads();
            for (int t = tid; t < c.ns * 2 * kModes; t += kThreads) {
                const int si = t / (2 * kModes), r = t % (2 * kModes), k = r % kModes;
                const CoStr s = c.str[si];
                const float* src = lds + (r < kModes ? 0 : c.tot_nodes) + s.node_off;
                const float* B = c.basis + (size_t)s.node_off * kModes + k;
                float acc = 0.0f;
#pragma unroll 8
                for (int i = 0; i < s.n_nodes; ++i) acc += B[(size_t)i * kModes] * src[i];
                c.modal[si * 2 * kModes + r] = acc * (2.0f / (float)(s.n_nodes > 1 ? s.n_nodes - 1 : 1));
            }
            __syncthreads();
        }
        if (a.sample_end) {
            const int S = c.tot_seg;
            float* v0 = lds;                       // [S] per-segment velocities
            float* v1 = lds + S;
            float* pr0 = lds + 2 * S;              // [S] products of one pickup row
            float* pr1 = lds + 3 * S;
            if (c.nmag > 0) {
                for (int si = 0; si < c.ns; ++si) {
                    const CoStr s = c.str[si];
                    for (int q = tid; q < s.n_nodes - 1; q += kThreads) {
                        const int n = s.node_off + q;
                        v0[s.seg_off + q] = (c.u[n] - c.u[2 * c.tot_nodes + n]) * s.inv_dt_s;
                        v1[s.seg_off + q] = (c.u[c.tot_nodes + n] - c.u[3 * c.tot_nodes + n]) * s.inv_dt_s;
                    }
                }
                __syncthreads();
                for (int r = 0; dfr && r < c.nmag; ++r) {  // products only, summed later
                    const float* k0 = c.K0 + (size_t)r * S;
                    const float* k1 = c.K1 + (size_t)r * S;
                    float* d = dprod + (size_t)r * 2 * S;
                    for (int j = tid; j < S; j += kThreads) {
                        d[j] = k0[j] * v0[j];
                        d[S + j] = k1[j] * v1[j];
                    }
                }
                for (int r = 0; !dfr && r < c.nmag; ++r) {  // :2152-2171
                    const float* k0 = c.K0 + (size_t)r * S;
                    const float* k1 = c.K1 + (size_t)r * S;
                    for (int j = tid; j < S; j += kThreads) {
                        pr0[j] = k0[j] * v0[j];
                        pr1[j] = k1[j] * v1[j];
                    }
                    __syncthreads();
                    if (tid == 0) ((float*)flag)[3] = seq_sum(pr0, S, 0.0f);
                    if (tid == 64) ((float*)flag)[4] = seq_sum(pr1, S, 0.0f);
                    __syncthreads();
                    if (tid == 0)
                        c.pk_out[(size_t)c.magmap[r] * kRing + a.out_pos] =
                            ((float*)flag)[3] + ((float*)flag)[4];
                    __syncthreads();
                }
            }
            const int nout = c.amr ? c.npk : c.npk + c.nmic;
            for (int t = tid; t < nout; t += kThreads) {
                if (t < c.npk) {
                    const int p = t;
                    if (c.pk_type[p] == PICKUP_PIEZO) {         // :1892-1904, :2174-2181
                        float acc = 0.0f;
                        for (int si = 0; si < c.ns; ++si) {
                            float v = 0.0f;
                            if (c.pk_mask[p] & (1u << (unsigned)si)) {
                                const CoStr s = c.str[si];
                                const int n = s.node_off + s.n_nodes;
                                const float s0 = (c.u[n - 1] - c.u[n - 2]) / s.ds;
                                const float s1 = (c.u[c.tot_nodes + n - 1] - c.u[c.tot_nodes + n - 2]) / s.ds;
                                v = c.pk_sens[p] * c.sw[si * kSw + kTeff] * sqrtf(s0 * s0 + s1 * s1);
                            }
                            acc += v;
                        }
                        c.pk_out[(size_t)p * kRing + a.out_pos] = acc;
                    } else if (c.nmag == 0) {
                        c.pk_out[(size_t)p * kRing + a.out_pos] = 0.0f;
                    }
                } else {                                        // :2183-2219
                    const int m = t - c.npk;
                    const int* id = c.mPi + 8 * m;
                    const float* wt = c.mPw + 8 * m;
                    const float P = wt[0] * a.P[id[0]] + wt[1] * a.P[id[1]] + wt[2] * a.P[id[2]] +
                                    wt[3] * a.P[id[3]] + wt[4] * a.P[id[4]] + wt[5] * a.P[id[5]] +
                                    wt[6] * a.P[id[6]] + wt[7] * a.P[id[7]];
                    float V[3] = {0.0f, 0.0f, 0.0f};
                    if (c.any_vel) {
                        for (int ax = 0; ax < 3; ++ax) {
                            const int* vi = c.mVi + (size_t)ax * c.nmic * 8 + 8 * m;
                            const float* vw = c.mVw + (size_t)ax * c.nmic * 8 + 8 * m;
                            float v = 0.0f;
                            for (int q = 0; q < 8; ++q) v += vw[q] * a.V[ax][vi[q]];
                            V[ax] = v;
                        }
                    }
                    const float* mp = c.mic_par + 6 * m;
                    float sig = P * mp[0];
                    if (mp[1] != 0.0f) {
                        const float vn = V[0] * mp[3] + V[1] * mp[4] + V[2] * mp[5];
                        sig += mp[1] * c.rhoc * vn;
                    }
                    c.mic_out[(size_t)m * kRing + a.out_pos] = sig * mp[2];
                }
            }
            if (c.amr && c.nmic > 0 && dfr) {
                // products of every mic's stencil entries, summed later
                float* pp = dprod + (size_t)2 * S * c.nmag;
                float* pv = pp + c.anp;
                const int p0 = c.amp_start[0], v0i = c.amv_start[0];
                for (int k = tid; k < c.anp; k += kThreads)
                    pp[k] = c.amp_w[p0 + k] * a.P[c.amp_cell[p0 + k]];
                for (int k = tid; k < c.anv; k += kThreads) {
                    const float vf = a.V[0][c.amv_face[v0i + k]];
                    pv[k] = c.amv_w[0][v0i + k] * vf;
                    pv[c.anv + k] = c.amv_w[1][v0i + k] * vf;
                    pv[2 * c.anv + k] = c.amv_w[2][v0i + k] * vf;
                }
            } else if (c.amr && c.nmic > 0) {
                // AMR mic stencils (acoustic_amr.cpp:1762-1792): the products of
                // a chunk of entries (all mics' entries are contiguous) are formed
                // in parallel, then each mic's running sums take its entries of
                // the chunk in stencil order, from +0
                float* pp = lds;                           // [kThreads]
                float* pv = lds + kThreads;                // [3][kThreads]
                float* acc = lds + 4 * kThreads;           // [4][nmic]: p, vx, vy, vz
                __syncthreads();
                for (int t = tid; t < 4 * c.nmic; t += kThreads) acc[t] = 0.0f;
                const int p0 = c.amp_start[0], p1 = c.amp_start[c.nmic];
                for (int base = p0; base < p1; base += kThreads) {
                    const int k = base + tid;
                    if (k < p1) pp[tid] = c.amp_w[k] * a.P[c.amp_cell[k]];
                    __syncthreads();
                    for (int m = tid; m < c.nmic; m += kThreads) {
                        const int lo = max(c.amp_start[m], base);
                        const int hi = min(c.amp_start[m + 1], base + kThreads);
                        float r = acc[m];
                        for (int j = lo; j < hi; ++j) r += pp[j - base];
                        acc[m] = r;
                    }
                    __syncthreads();
                }
                const int v0 = c.amv_start[0], v1 = c.amv_start[c.nmic];
                for (int base = v0; base < v1; base += kThreads) {
                    const int k = base + tid;
                    if (k < v1) {
                        const float vf = a.V[0][c.amv_face[k]];
                        pv[tid] = c.amv_w[0][k] * vf;
                        pv[kThreads + tid] = c.amv_w[1][k] * vf;
                        pv[2 * kThreads + tid] = c.amv_w[2][k] * vf;
                    }
                    __syncthreads();
                    for (int t = tid; t < 3 * c.nmic; t += kThreads) {
                        const int ax = t / c.nmic, m = t - ax * c.nmic;
                        const int lo = max(c.amv_start[m], base);
                        const int hi = min(c.amv_start[m + 1], base + kThreads);
                        float r = acc[(1 + ax) * c.nmic + m];
                        for (int j = lo; j < hi; ++j) r += pv[ax * kThreads + j - base];
                        acc[(1 + ax) * c.nmic + m] = r;
                    }
                    __syncthreads();
                }
                __syncthreads();
                for (int m = tid; m < c.nmic; m += kThreads) {     // :2208-2219
                    const float* mp = c.mic_par + 6 * m;
                    float sig = acc[m] * mp[0];
                    if (mp[1] != 0.0f) {
                        const float vn = acc[c.nmic + m] * mp[3] + acc[2 * c.nmic + m] * mp[4] +
                                         acc[3 * c.nmic + m] * mp[5];
                        sig += mp[1] * c.rhoc * vn;
                    }
                    c.mic_out[(size_t)m * kRing + a.out_pos] = sig * mp[2];
                }
            }
        }
    }
    __syncthreads();
    if (!a.run) return;

    // The plate update of the previous step consumed the force accumulator;
    // its clear was left to this kernel (acoustic_fdtd.cpp:1047-1049).
    if (a.post && c.clear_ext)
        for (int p = tid; p < c.Nplate; p += kThreads) a.ext[p] = 0.0f;

    // ── neck mode (:1934-1983) on the last wavefront, beside the strings'
    // loads and first tension (the two do not depend on each other) ──
    constexpr int SW = kThreads / 64 - 1;      // string wavefronts
    // The neck plate sum over the current plate: formed by the previous body
    // step's sweep (same products, same order) when it ran that side job and
    // its tag names that step, else here.
    const bool pre = c.neck_on && a.npre && ((const int*)c.npre)[1] == a.step - 1;
    if (c.neck_on) {
        // out-of-range plate indices are skipped by the reference; adding +0
        // to the running sum (which starts at +0, so never holds -0) is the same
        for (int i = tid; i < c.nn && !pre; i += kThreads) {
            const int p = c.nidx[i];
            nprod[i] = (p < 0 || p >= c.Nplate) ? 0.0f : a.w[p] * c.nwgt[i];
        }
        for (int si = tid; si < c.ns; si += kThreads) {
            fsT[si] = c.sw[si * kSw + kTeff];
            fsU[si] = c.u[c.str[si].node_off + 1];
        }
        __syncthreads();
    }
    if (wave == SW) {
        if (c.neck_on && lane == 0) {
            const float pw = pre ? c.npre[0] : seq_sum(nprod, c.nn, 0.0f);
            const float nw0 = c.neck[0], nwp = c.neck[1], npw = c.neck[2];
            float fs = 0.0f;
            for (int si = 0; si < c.ns; ++si) {
                const CoStr s = c.str[si];
                if (s.n_nodes > 1) fs += fsT[si] * (fsU[si] - nw0) / s.ds;
            }
            const float idt = 1.0f / c.dt_b, idt2 = idt * idt;
            const float pv = (pw - npw) * idt;
            const float m = c.n_m, cc = c.n_c, k = c.n_k;
            const float den = m * idt2 + 0.5f * cc * idt;
            const float rhs = fs + c.nb_k * pw + c.nb_c * pv + (2.0f * m * idt2 - k) * nw0 +
                              (-m * idt2 + 0.5f * cc * idt) * nwp;
            const float nw = (den > 1e-12f) ? rhs / den : pw;
            float force = 0.0f;
            bool bad = !__builtin_isfinite(nw) || fabsf(nw) > 0.05f;
            if (!bad) {
                const float nv = (nw - nw0) * idt;
                force = c.nb_k * (nw - pw) + c.nb_c * (nv - pv);
                bad = !__builtin_isfinite(force);
                c.neck[1] = nw0;
                c.neck[0] = nw;
                if (!bad) c.neck[2] = pw;
            }
            if (bad) fail_here(c, a, &flag[0]);
            ((float*)flag)[1] = nw;
            ((float*)flag)[2] = force;
        }
    } else {
        // string loads, the direct drive and the tension of the first sub-step
        for (int si = wave; si < c.ns; si += SW) {
            const CoStr s = c.str[si];
            const int N = s.n_nodes;
            // this launch's copies of the string's state words and force
            // (the saddle word comes from the early gather below when it runs)
            if (lane < kSw && !(lane == kSad && c.sad_early && a.post))
                swl[si * kSw + lane] = c.sw[si * kSw + lane];
            if (lane == kSw) psub[si] = c.blk_sub[a.sample * c.ns + si];
            if (lane == kSw + 1) pnode[si] = c.blk_node[a.sample * c.ns + si];
            float* base = lds + 8 * s.node_off;
            float* fe0 = base + 6 * N;
            float* scr = base + 7 * N;
            for (int i = lane; i < N; i += 64) {
                const int n = s.node_off + i;
                base[i] = c.u[n];
                base[N + i] = c.u[2 * c.tot_nodes + n];
                base[3 * N + i] = c.u[c.tot_nodes + n];
                base[4 * N + i] = c.u[3 * c.tot_nodes + n];
                fe0[i] = 0.0f;
            }
            wave_sync();
            if (lane == 0 && a.sub == 0 && c.blk_dval) {          // direct drive (:2300-2309)
                const float d = c.blk_dval[a.sample * c.ns + si];
                if (d != 0.0f) fe0[c.blk_dnode[si]] += d;
            }
            // the shape's arc length is the last shape check's (kArcOk), or is
            // summed here (first sub-step after creation or reset)
            bool tb;
            if (c.sw[si * kSw + kArcOk] != 0.0f) {
                int b = 0;
                if (lane == 0) b = tension_arc(s, c.sw[si * kSw + kArc], &teff0[si]);
                tb = __shfl(b, 0) != 0;
            } else {
                tb = tension_of(base, base + 3 * N, scr, N, s, lane, &teff0[si]);
            }
            if (tb && lane == 0) fail_here(c, a, &flag[0]);
        }
        // wavefronts without a string: the bridge kernels into LDS, and (n_strings
        // <= 12) the previous sub-step's saddle gather (:2105-2125), which only
        // the bridge phase reads
        if (c.sbk_lds && wave >= c.ns)
            for (int i = (wave - c.ns) * 64 + lane; i < c.ns * c.nb; i += (SW - c.ns) * 64)
                sbkl[i] = c.sbk[i];
        if (c.sad_early && a.post && wave == SW - 3) {
            for (int si = lane; si < c.ns; si += 64) {
                const int* id = c.sad_idx + 4 * si;
                const float* g = c.sad_wgt + 4 * si;
                const float t0 = c.amr ? 0.0f + g[0] * a.w[id[0]] : g[0] * a.w[id[0]];
                const float sad = t0 + g[1] * a.w[id[1]] + g[2] * a.w[id[2]] + g[3] * a.w[id[3]];
                c.sw[si * kSw + kSad] = sad;
                swl[si * kSw + kSad] = sad;
            }
        }
        if (a.sample_end && c.defer && wave == SW - 1 && c.nmag > 0) {
            // pickup rows (:2152-2171): lane 2r sums row r's K0 products, lane
            // 2r + 1 its K1 products, each in index order
            const int S = c.tot_seg;
            float sum = 0.0f;
            if (lane < 2 * c.nmag)
                sum = seq_sum(dprod + (size_t)(lane >> 1) * 2 * S + (lane & 1) * S, S, 0.0f);
            const float other = __shfl_xor(sum, 1);
            if (lane < 2 * c.nmag && (lane & 1) == 0)
                c.pk_out[(size_t)c.magmap[lane >> 1] * kRing + a.out_pos] = sum + other;
        }
        if (a.sample_end && c.defer && wave == SW - 2 && c.amr && c.nmic > 0) {
            // AMR mic stencil sums (acoustic_amr.cpp:1762-1792), one lane per
            // mic and component, then the polar mix (:2208-2219)
            const int S = c.tot_seg, M = c.nmic;
            const float* pp = dprod + (size_t)2 * S * c.nmag;
            const float* pv = pp + c.anp;
            float* dres = (float*)pv + 3 * c.anv;
            if (lane < 4 * M) {
                const int kind = lane / M, m = lane - kind * M;
                float r;
                if (kind == 0) {
                    const int lo = c.amp_start[m] - c.amp_start[0];
                    r = seq_sum(pp + lo, c.amp_start[m + 1] - c.amp_start[m], 0.0f);
                } else {
                    const int lo = c.amv_start[m] - c.amv_start[0];
                    r = seq_sum(pv + (size_t)(kind - 1) * c.anv + lo,
                                c.amv_start[m + 1] - c.amv_start[m], 0.0f);
                }
                dres[lane] = r;
            }
            wave_sync();
            if (lane < M) {
                const float* mp = c.mic_par + 6 * lane;
                float sig = dres[lane] * mp[0];
                if (mp[1] != 0.0f) {
                    const float vn = dres[M + lane] * mp[3] + dres[2 * M + lane] * mp[4] +
                                     dres[3 * M + lane] * mp[5];
                    sig += mp[1] * c.rhoc * vn;
                }
                c.mic_out[(size_t)lane * kRing + a.out_pos] = sig * mp[2];
            }
        }
    }
    __syncthreads();
    if (flag[0]) return;
    const float nut = c.neck_on ? ((float*)flag)[1] : 0.0f;

    // ── strings (:1987-2003): one wavefront per string ──
    for (int si = wave; si < c.ns && wave < SW; si += SW) {
        const CoStr s = c.str[si];
        const int N = s.n_nodes;
        float* base = lds + 8 * s.node_off;
        float* buf[6];
        for (int b = 0; b < 6; ++b) buf[b] = base + b * N;
        float* fe0 = base + 6 * N;
        float* scr = base + 7 * N;
        int r0u = 0, r0p = 1, r0t = 2, r1u = 3, r1p = 4, r1t = 5;
        float Teff = teff0[si];
        const float bridge_w = swl[si * kSw + kBr];
        const float per_sub = psub[si];
        const int fnode = pnode[si];
        bool failed = false;
        float arc = 0.0f;                       // lane 0: arc length of the checked shape
        for (int q = 0; q < s.n_sub && !failed; ++q) {
            // geometric tension (:563-593); the first sub-step's was taken above,
            // later ones take the previous shape check's arc length
            if (q > 0) {
                int b = 0;
                if (lane == 0) b = tension_arc(s, arc, &teff0[si]);
                if (__shfl(b, 0)) {
                    failed = true;
                    break;
                }
            }
            wave_sync();
            Teff = teff0[si];
            if (lane == 0 && per_sub != 0.0f) fe0[fnode] += per_sub;
            wave_sync();
            const float cfl = sqrtf(Teff / s.mu) * s.dt_s / s.ds;
            const float c2 = cfl * cfl;
            const float h = 0.5f * s.gamma_eff * s.dt_s;
            const float inv = 1.0f / (1.0f + h);
            const float omh = 1.0f - h;
            if (string_pol(buf[r0u], buf[r0p], buf[r0t], fe0, N, c2, omh, inv, s.biharm, bridge_w,
                           nut, lane)) { failed = true; break; }
            for (int i = lane; i < N; i += 64) fe0[i] = 0.0f;
            { const int t = r0p; r0p = r0u; r0u = r0t; r0t = t; }
            wave_sync();
            if (string_pol(buf[r1u], buf[r1p], buf[r1t], nullptr, N, c2, omh, inv, s.biharm, 0.0f,
                           0.0f, lane)) { failed = true; break; }
            { const int t = r1p; r1p = r1u; r1u = r1t; r1t = t; }
            wave_sync();
            // shape validity (:538-561)
            bool bad = false;
            for (int i = lane; i < N - 1; i += 64) {
                const float d0 = buf[r0u][i + 1] - buf[r0u][i];
                const float d1 = buf[r1u][i + 1] - buf[r1u][i];
                if (!__builtin_isfinite(d0) || !__builtin_isfinite(d1)) bad = true;
                const float sl = sqrtf(d0 * d0 + d1 * d1) / s.ds;
                if (!__builtin_isfinite(sl) || sl > kMaxSlope) bad = true;
                scr[i] = sqrtf(s.ds * s.ds + d0 * d0 + d1 * d1);
            }
            wave_sync();
            if (wave_any(bad)) { failed = true; break; }
            int sbad = 0;
            if (lane == 0) {
                arc = seq_sum(scr, N - 1, 0.0f);
                const float fr = (arc - s.L) / s.L;
                if (!__builtin_isfinite(fr) || fr > kMaxStretch) sbad = 1;
            }
            if (__shfl(sbad, 0)) { failed = true; break; }
            wave_sync();
        }
        if (failed) {
            if (lane == 0) fail_here(c, a, &flag[0]);
            break;
        }
        for (int i = lane; i < N; i += 64) {
            const int n = s.node_off + i;
            c.u[n] = buf[r0u][i];
            c.u[2 * c.tot_nodes + n] = buf[r0p][i];
            c.u[c.tot_nodes + n] = buf[r1u][i];
            c.u[3 * c.tot_nodes + n] = buf[r1p][i];
        }
        if (lane == 0) {
            c.sw[si * kSw + kTeff] = Teff;
            c.sw[si * kSw + kArc] = arc;
            c.sw[si * kSw + kArcOk] = 1.0f;
            fsU[si] = buf[r0u][N - 2];                 // for the bridge below
            fsT[si] = Teff;
        }
    }
    __syncthreads();
    if (flag[0]) return;

    // ── bridge impedances and bridge drive (:2009-2062) ──
    if (c.nb > 0 && c.ns > 0) {
        if (tid < c.ns) {
            const CoStr s = c.str[tid];
            float* w = c.sw + tid * kSw;
            const float* wl = swl + tid * kSw;
            const float br = wl[kBr], brp = wl[kBrP], sad = wl[kSad], sadp = wl[kSadP];
            const float sl = (fsU[tid] - br) / s.ds;
            bool bad = !__builtin_isfinite(sl) || fabsf(sl) > 1.0f;
            if (!bad) {
                const float fstr = fsT[tid] * sl;
                const float m = s.b_m, cc = s.b_c, k = s.b_k;
                const float idt = 1.0f / c.dt_b, idt2 = idt * idt;
                const float pv = (sad - sadp) * idt;
                const float den = m * idt2 + 0.5f * cc * idt;
                const float rhs = fstr + k * sad + cc * pv + (2.0f * m * idt2 - k) * br +
                                  (-m * idt2 + 0.5f * cc * idt) * brp;
                const float bn = (den > 1e-12f) ? rhs / den : sad;
                bad = !__builtin_isfinite(bn) || fabsf(bn) > 0.05f;
                if (!bad) {
                    w[kBrP] = br;
                    w[kBr] = bn;
                    const float bv = (bn - br) * idt;
                    const float fp = k * (bn - sad) + cc * (bv - pv);
                    bad = !__builtin_isfinite(fp);
                    fpl[tid] = fp;
                    if (!bad) w[kSadP] = sad;
                }
            }
            if (bad) fail_here(c, a, &flag[0]);
        }
        __syncthreads();
        if (flag[0]) return;
        for (int cix = tid; cix < c.nb; cix += kThreads) {
            float acc = 0.0f;
            const float* kb = c.sbk_lds ? sbkl : c.sbk;
            for (int si = 0; si < c.ns; ++si) acc += kb[(size_t)si * c.nb + cix] * fpl[si];
            drive[cix] = acc;
        }
        __syncthreads();
        // uniform fdtd_inject_bridge_drive(drive, 1): ext[p] += 1 * d * inv_dx2 per
        // entry; AMR amr_inject_bridge_drive: ext[p] += d * w * 1; in order
        for (int g = tid; g < c.nbgrp; g += kThreads) {
            const int* start = c.bgrp;
            const int* node = c.bgrp + c.nbgrp + 1;
            const int* ent = c.bgrp + 2 * c.nbgrp + 1;
            const int p = node[g];
            float acc = a.ext[p];
            if (c.amr)
                for (int q = start[g]; q < start[g + 1]; ++q)
                    acc += drive[ent[q]] * c.bwgt[ent[q]] * 1.0f;
            else
                for (int q = start[g]; q < start[g + 1]; ++q) acc += 1.0f * drive[ent[q]] * c.inv_dx2;
            a.ext[p] = acc;
        }
        __syncthreads();
    }
    // ── neck reaction into the plate (fdtd_inject_plate_force, :2085-2098) ──
    if (c.neck_on) {
        const float total = ((float*)flag)[2];
        if (total != 0.0f) {
            for (int g = tid; g < c.nngrp; g += kThreads) {
                const int* start = c.ngrp;
                const int* node = c.ngrp + c.nngrp + 1;
                const int* ent = c.ngrp + 2 * c.nngrp + 1;
                const int p = node[g];
                float acc = a.ext[p];
                // uniform: total * w / dx^2; AMR amr_inject_plate_force: w * total
                if (c.amr)
                    for (int q = start[g]; q < start[g + 1]; ++q) acc += total * c.nwgt[ent[q]];
                else
                    for (int q = start[g]; q < start[g + 1]; ++q)
                        acc += total * c.nwgt[ent[q]] * c.inv_dx2;
                a.ext[p] = acc;
            }
        }
    }
}

// ── host side ───────────────────────────────────────────────────────────

struct HString {
    int n_segs, n_nodes, n_sub, node_off, seg_off;
    std::vector<float> xyz, tan, nrm, bin;
    float L, ds, T0, EA, mu, c_s, gamma, dt_s, biharm;
    float b_m, b_k, b_c;
    std::vector<float> fbuf;
    int fnode, fonset;
};

struct HPluck {
    int onset, si;
    float pos, amp;
    int ndur;
};

template <typename T>
hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

template <typename T>
hipError_t dupload(T** p, const std::vector<T>& v) {
    hipError_t e = dalloc(p, v.size());
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    return hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

// Grouped force entries by target plate node, stable in call order (the
// layout of the library's force accumulation: start | node | entry).
int group_entries(const int* idx, int n, int div, const uint8_t* mask, int nmax,
                  std::vector<int>& grp) {
    std::vector<std::pair<int, int>> e;
    for (int i = 0; i < n; ++i) {
        const int p = idx[i] / div;
        if (p < 0 || p >= nmax) continue;
        if (mask && !mask[p]) continue;
        e.emplace_back(p, i);
    }
    std::stable_sort(e.begin(), e.end(),
                     [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
                         return x.first < y.first;
                     });
    std::vector<int> start, node;
    for (size_t q = 0; q < e.size(); ++q)
        if (q == 0 || e[q].first != e[q - 1].first) {
            start.push_back((int)q);
            node.push_back(e[q].first);
        }
    start.push_back((int)e.size());
    grp.clear();
    grp.insert(grp.end(), start.begin(), start.end());
    grp.insert(grp.end(), node.begin(), node.end());
    for (const auto& x : e) grp.push_back(x.second);
    return (int)node.size();
}

}  // namespace

struct AcousticCoEvolverState {
    AcousticFDTDState* fdtd = nullptr;     // uniform body, or
    AcousticAMRState* amr = nullptr;       // AMR body (coevolver_create_amr)
    int ns = 0, npk = 0, nmic = 0, Nx = 0, Ny = 0, Nz = 0, plate_iz = 0, K = 1;
    float dx = 0, origin[3] = {0, 0, 0}, rho_air = 0, c = 0, sr = 0, dt_audio = 0, dt_fdtd = 0;
    float gamma_scale = 1.0f, gamma_dev = -1.0f;
    int modal_stride = 1, modal_counter = 0, out_pos = 0, written = 0, cur_sample = 0;
    std::vector<HString> s;
    std::vector<int> pk_type;
    std::vector<unsigned> pk_mask;
    std::vector<float> pk_sens;
    std::vector<float> mic_g;               // grid coordinates [nmic][3]
    int any_vel = 0, nmag = 0, tot_seg = 0, tot_nodes = 0;
    int nb = 0, nn = 0;
    std::vector<int> bidx, nidx;
    std::vector<float> bwgt, nwgt;
    float n_m = 0, n_k = 0, n_c = 0, nb_k = 0, nb_c = 0;
    std::vector<HPluck> ev;
    int cursor = 0;
    size_t lds_bytes = 0;
    CoConst k{};
    // device buffers
    CoStr* d_str = nullptr;
    float *d_sw = nullptr, *d_u = nullptr, *d_neck = nullptr, *d_nwgt = nullptr;
    int *d_nidx = nullptr, *d_ngrp = nullptr;
    float* d_sbk = nullptr;
    float* d_bwgt = nullptr;               // AMR body: bridge plate weights
    float* d_npre = nullptr;               // uniform body: neck plate sum from the sweep
    int* d_bgrp = nullptr;                 // AMR body: bridge force groups
    int* d_sad_idx = nullptr;
    float* d_sad_wgt = nullptr;
    float *d_basis = nullptr, *d_modal = nullptr, *d_K0 = nullptr, *d_K1 = nullptr;
    int *d_magmap = nullptr, *d_pk_type = nullptr;
    unsigned* d_pk_mask = nullptr;
    float* d_pk_sens = nullptr;
    int *d_mPi = nullptr, *d_mVi = nullptr;
    float *d_mPw = nullptr, *d_mVw = nullptr, *d_mic_par = nullptr;
    float *d_pk_out = nullptr, *d_mic_out = nullptr;
    float *d_blk_sub = nullptr, *d_blk_dval = nullptr;
    int *d_blk_node = nullptr, *d_blk_dnode = nullptr;
    int blk_cap = 0;
    // async worker (acoustic_coevolver.cpp:155-162)
    std::thread worker;
    std::atomic<int> status{CE_STATUS_IDLE}, progress{0}, total{0}, cancel{0}, last_error{CE_OK};
};

namespace {

bool busy(const AcousticCoEvolverState* st) {
    return st && st->status.load(std::memory_order_acquire) == CE_STATUS_RUNNING;
}

void co_free(AcousticCoEvolverState* st) {
    void* ptrs[] = {st->d_str, st->d_sw, st->d_u, st->d_neck, st->d_nwgt, st->d_nidx,
                    st->d_ngrp, st->d_sbk, st->d_sad_idx, st->d_sad_wgt, st->d_basis,
                    st->d_modal, st->d_K0, st->d_K1, st->d_magmap, st->d_pk_type,
                    st->d_pk_mask, st->d_pk_sens, st->d_mPi, st->d_mVi, st->d_mPw, st->d_mVw,
                    st->d_mic_par, st->d_pk_out, st->d_mic_out, st->d_blk_sub, st->d_blk_dval,
                    st->d_blk_node, st->d_blk_dnode, st->d_bwgt, st->d_bgrp, st->d_npre};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (st->fdtd) fdtd_destroy(st->fdtd);
    st->fdtd = nullptr;
    if (st->amr) amr_destroy(st->amr);
    st->amr = nullptr;
}

// Frame (parallel transport from a world-Y or world-Z bootstrap) —
// acoustic_coevolver.cpp:325-396.
void string_frame(HString& s) {
    const int N = s.n_segs;
    s.tan.assign(3 * N, 0.0f);
    s.nrm.assign(3 * N, 0.0f);
    s.bin.assign(3 * N, 0.0f);
    s.L = 0.0f;
    for (int q = 0; q < N; ++q) {
        const float dx = s.xyz[(q + 1) * 3 + 0] - s.xyz[q * 3 + 0];
        const float dy = s.xyz[(q + 1) * 3 + 1] - s.xyz[q * 3 + 1];
        const float dz = s.xyz[(q + 1) * 3 + 2] - s.xyz[q * 3 + 2];
        float len = sqrtf(dx * dx + dy * dy + dz * dz);
        if (len < 1e-12f) len = 1e-12f;
        s.L += len;
        s.tan[q * 3 + 0] = dx / len;
        s.tan[q * 3 + 1] = dy / len;
        s.tan[q * 3 + 2] = dz / len;
    }
    s.ds = s.L / (float)N;
    const float ty = s.tan[1];
    float ref[3] = {0.0f, 0.0f, 0.0f};
    if (fabsf(ty) < 0.9f) ref[1] = 1.0f; else ref[2] = 1.0f;
    const float tx = s.tan[0], tz = s.tan[2];
    float dot = ref[0] * tx + ref[1] * ty + ref[2] * tz;
    float nx = ref[0] - dot * tx, ny = ref[1] - dot * ty, nz = ref[2] - dot * tz;
    float nl = sqrtf(nx * nx + ny * ny + nz * nz);
    if (nl < 1e-12f) nl = 1e-12f;
    nx /= nl; ny /= nl; nz /= nl;
    for (int q = 0; q < N; ++q) {
        const float* t = &s.tan[q * 3];
        if (q > 0) {
            const float* tp = &s.tan[(q - 1) * 3];
            float ax = tp[1] * t[2] - tp[2] * t[1];
            float ay = tp[2] * t[0] - tp[0] * t[2];
            float az = tp[0] * t[1] - tp[1] * t[0];
            const float al = sqrtf(ax * ax + ay * ay + az * az);
            if (al > 1e-8f) {
                const float th = asinf(fminf(al, 1.0f));
                const float c_ = cosf(th), s_ = sinf(th) / al;
                ax *= s_; ay *= s_; az *= s_;
                const float cx = ay * nz - az * ny, cy = az * nx - ax * nz, cz = ax * ny - ay * nx;
                const float adn = ax * nx + ay * ny + az * nz;
                nx = c_ * nx + sinf(th) * cx + (1 - c_) * adn * ax;
                ny = c_ * ny + sinf(th) * cy + (1 - c_) * adn * ay;
                nz = c_ * nz + sinf(th) * cz + (1 - c_) * adn * az;
                dot = nx * t[0] + ny * t[1] + nz * t[2];
                nx -= dot * t[0]; ny -= dot * t[1]; nz -= dot * t[2];
                nl = sqrtf(nx * nx + ny * ny + nz * nz);
                if (nl < 1e-12f) nl = 1e-12f;
                nx /= nl; ny /= nl; nz /= nl;
            }
        }
        float* n = &s.nrm[q * 3];
        float* b = &s.bin[q * 3];
        n[0] = nx; n[1] = ny; n[2] = nz;
        b[0] = t[1] * n[2] - t[2] * n[1];
        b[1] = t[2] * n[0] - t[0] * n[2];
        b[2] = t[0] * n[1] - t[1] * n[0];
    }
}

// String parameters — acoustic_coevolver.cpp:399-522.
bool string_init(HString& s, const CoEvolverStringDef& d, float dt_fdtd) {
    const int N = d.n_segs;
    if (N < 2 || !d.path_xyz || d.tension_N <= 0 || d.linear_mass_kgm <= 0) return false;
    s.n_segs = N;
    s.n_nodes = N + 1;
    s.xyz.assign(d.path_xyz, d.path_xyz + 3 * (N + 1));
    s.T0 = d.tension_N;
    s.mu = d.linear_mass_kgm;
    s.gamma = d.damping;
    s.EA = d.axial_stiffness_N;
    if (s.EA <= 0.0f && s.mu > 0.0f) s.EA = (2.0e11f / 7850.0f) * s.mu;
    s.b_m = 0.00035f;
    const float ob = 2.0f * kPi * 450.0f;
    s.b_k = s.b_m * ob * ob;
    s.b_c = s.b_m * ob / 2.5f;
    string_frame(s);
    s.c_s = sqrtf(d.tension_N / d.linear_mass_kgm);
    const float EI = d.stiffness_EI, mu = d.linear_mass_kgm;
    const float cst = (EI > 0.0f && mu > 0.0f) ? (EI / mu) * (kPi / s.ds) * (kPi / s.ds) : 0.0f;
    const float c_eff = sqrtf(s.c_s * s.c_s + cst);
    const float dtmax = 0.95f * s.ds / c_eff;
    s.n_sub = (int)ceilf(dt_fdtd / dtmax);
    if (s.n_sub < 1) s.n_sub = 1;
    s.dt_s = dt_fdtd / (float)s.n_sub;
    s.biharm = (EI > 0.0f && mu > 0.0f)
                   ? (EI / mu) * (s.dt_s * s.dt_s) / (s.ds * s.ds * s.ds * s.ds) : 0.0f;
    s.fnode = 0;
    s.fonset = 0;
    return true;
}

float line_dist(const float* p, const float* a, const float* d) {
    const float ax = p[0] - a[0], ay = p[1] - a[1], az = p[2] - a[2];
    const float pr = ax * d[0] + ay * d[1] + az * d[2];
    const float rx = ax - pr * d[0], ry = ay - pr * d[1], rz = az - pr * d[2];
    return sqrtf(rx * rx + ry * ry + rz * rz);
}

// Magnetic B-kernels projected on the segment frame — :711-775.
void bkernels(const CoEvolverPickupDef& d, const HString& s, std::vector<float>& Bn,
              std::vector<float>& Bb) {
    float ax = d.axis[0], ay = d.axis[1], az = d.axis[2];
    const float al = sqrtf(ax * ax + ay * ay + az * az);
    if (al < 1e-9f) { ax = 0.0f; ay = 1.0f; az = 0.0f; }
    else { ax /= al; ay /= al; az /= al; }
    const float i2s2 = (d.pole_sigma > 0.0f) ? 0.5f / (d.pole_sigma * d.pole_sigma) : 1e12f;
    const float sp = d.coil_spacing;
    const float au[3] = {ax, ay, az};
    Bn.assign(s.n_segs, 0.0f);
    Bb.assign(s.n_segs, 0.0f);
    for (int q = 0; q < s.n_segs; ++q) {
        float m[3];
        for (int k = 0; k < 3; ++k) m[k] = 0.5f * (s.xyz[q * 3 + k] + s.xyz[(q + 1) * 3 + k]);
        float B;
        if (d.type == PICKUP_SINGLE_COIL) {
            const float r = line_dist(m, d.pos, au);
            B = expf(-r * r * i2s2);
        } else if (d.type == PICKUP_HUMBUCKER) {
            const float p1[3] = {d.pos[0] + 0.5f * sp * ax, d.pos[1] + 0.5f * sp * ay,
                                 d.pos[2] + 0.5f * sp * az};
            const float p2[3] = {d.pos[0] - 0.5f * sp * ax, d.pos[1] - 0.5f * sp * ay,
                                 d.pos[2] - 0.5f * sp * az};
            const float r1 = line_dist(m, p1, au), r2 = line_dist(m, p2, au);
            B = expf(-r1 * r1 * i2s2) - expf(-r2 * r2 * i2s2);
        } else {
            B = 0.0f;
        }
        const float* n = &s.nrm[q * 3];
        const float* b = &s.bin[q * 3];
        Bn[q] = B * (n[0] * ax + n[1] * ay + n[2] * az);
        Bb[q] = B * (b[0] * ax + b[1] * ay + b[2] * az);
    }
}

int flat_clamped(const AcousticCoEvolverState* st, int i, int j, int k) {   // :786-793
    if (i < 0) i = 0;
    if (i >= st->Nx) i = st->Nx - 1;
    if (j < 0) j = 0;
    if (j >= st->Ny) j = st->Ny - 1;
    if (k < 0) k = 0;
    if (k >= st->Nz) k = st->Nz - 1;
    return k + st->Nz * (j + st->Ny * i);
}

// Bridge injection cells within 3 sigma of the saddles, one layer below the
// plate — :799-861.
void bridge_cells(AcousticCoEvolverState* st, int n, const float* xyz, float sigma) {
    int pz = st->plate_iz - 1;
    if (pz < 0) pz = 0;
    const int rc = (int)ceilf(3.0f * sigma / st->dx) + 1;
    int mx = n * (2 * rc + 1) * (2 * rc + 1);
    if (mx < 1) mx = 1;
    std::vector<int> idx;
    std::vector<float> wgt;
    const float i2s2 = 0.5f / (sigma * sigma);
    for (int q = 0; q < n; ++q) {
        const float gx = (xyz[q * 3 + 0] - st->origin[0]) / st->dx;
        const float gy = (xyz[q * 3 + 1] - st->origin[1]) / st->dx;
        const int ci0 = (int)roundf(gx), cj0 = (int)roundf(gy);
        for (int di = -rc; di <= rc; ++di)
            for (int dj = -rc; dj <= rc; ++dj) {
                const int ci = ci0 + di, cj = cj0 + dj;
                if (ci < 0 || ci >= st->Nx || cj < 0 || cj >= st->Ny) continue;
                const float wx = (float)ci * st->dx + st->origin[0] - xyz[q * 3 + 0];
                const float wy = (float)cj * st->dx + st->origin[1] - xyz[q * 3 + 1];
                const float w = expf(-(wx * wx + wy * wy) * i2s2);
                if (w < 1e-4f) continue;
                const int flat = flat_clamped(st, ci, cj, pz);
                bool dup = false;
                for (size_t e = 0; e < idx.size(); ++e)
                    if (idx[e] == flat) { wgt[e] += w; dup = true; break; }
                if (!dup && (int)idx.size() < mx) { idx.push_back(flat); wgt.push_back(w); }
            }
    }
    float ws = 0.0f;
    for (float w : wgt) ws += w;
    if (ws > 1e-12f)
        for (float& w : wgt) w /= ws;
    st->bidx = idx;
    st->bwgt = wgt;
    st->nb = (int)idx.size();
}

// Neck-joint plate kernel around the highest active plate row — :868-940.
int neck_cells(AcousticCoEvolverState* st, const uint8_t* pa, float sigma) {
    if (sigma <= 0.0f) sigma = 0.030f;
    float ymax = -FLT_MAX;
    for (int i = 0; i < st->Nx; ++i)
        for (int j = 0; j < st->Ny; ++j) {
            if (!pa[j + st->Ny * i]) continue;
            const float y = (float)j * st->dx + st->origin[1];
            if (y > ymax) ymax = y;
        }
    if (ymax == -FLT_MAX) return CE_ERR_PARAM;
    const float cx = 0.0f, cy = ymax - 0.012f;
    const int rc = (int)ceilf(3.0f * sigma / st->dx) + 1;
    const int ci = (int)floorf((cx - st->origin[0]) / st->dx);
    const int cj = (int)floorf((cy - st->origin[1]) / st->dx);
    const float i2s2 = 0.5f / (sigma * sigma);
    std::vector<int> idx;
    std::vector<float> wgt;
    float ws = 0.0f;
    for (int di = -rc; di <= rc; ++di) {
        const int i = ci + di;
        if (i < 0 || i >= st->Nx) continue;
        for (int dj = -rc; dj <= rc; ++dj) {
            const int j = cj + dj;
            if (j < 0 || j >= st->Ny) continue;
            const int p = j + st->Ny * i;
            if (!pa[p]) continue;
            const float x = (float)i * st->dx + st->origin[0];
            const float y = (float)j * st->dx + st->origin[1];
            const float ddx = x - cx, ddy = y - cy;
            const float ww = expf(-(ddx * ddx + ddy * ddy) * i2s2);
            if (ww <= 1e-12f) continue;
            idx.push_back(p);
            wgt.push_back(ww);
            ws += ww;
        }
    }
    if (idx.empty() || ws <= 1e-12f) return CE_ERR_PARAM;
    for (float& w : wgt) w /= ws;
    st->nidx = idx;
    st->nwgt = wgt;
    st->nn = (int)idx.size();
    return CE_OK;
}

// Plate mask with the canonical soundhole (centre (0, 0.050) m, radius
// 0.028 m) carved when the grid contains it — :946-981.
std::vector<uint8_t> carve_soundhole(const CoEvolverBodyDef* b) {
    const int Np = b->Nx * b->Ny;
    std::vector<uint8_t> out(b->plate_active, b->plate_active + Np);
    const float cx = 0.0f, cy = 0.050f, r = 0.028f, r2 = r * r;
    const float x0 = b->origin[0], y0 = b->origin[1];
    const float x1 = x0 + (float)(b->Nx - 1) * b->dx, y1 = y0 + (float)(b->Ny - 1) * b->dx;
    if (cx + r < x0 || cx - r > x1 || cy + r < y0 || cy - r > y1) return out;
    for (int i = 0; i < b->Nx; ++i) {
        const float x = b->origin[0] + (float)i * b->dx;
        for (int j = 0; j < b->Ny; ++j) {
            const float y = b->origin[1] + (float)j * b->dx;
            const float ddx = x - cx, ddy = y - cy;
            if (ddx * ddx + ddy * ddy < r2) out[j + b->Ny * i] = 0;
        }
    }
    return out;
}

// Padded-layout index of a packed staggered velocity index (the body FDTD
// keeps all three components on the cell grid).
int packed_to_padded_v(int ax, int q, int Nx, int Ny, int Nz) {
    (void)Nx;
    if (ax == 0) return q;
    if (ax == 1) {
        const int k = q % Nz, t = q / Nz, j = t % (Ny - 1), i = t / (Ny - 1);
        return k + Nz * (j + Ny * i);
    }
    const int k = q % (Nz - 1), t = q / (Nz - 1), j = t % Ny, i = t / Ny;
    return k + Nz * (j + Ny * i);
}

// ── per-block host preparation (the reference's per-sample scalar work) ──

struct BlockPlan {
    std::vector<float> sub;     // [n][ns]
    std::vector<int> node;      // [n][ns]
    std::vector<float> dval;    // [n][ns] (block-drive form)
    std::vector<int> dnode;     // [ns]
};

// Pluck ramp into a string's timed force buffer — :2413-2470.
int pluck_host(AcousticCoEvolverState* st, int si, float pos_n, float amp_in, int ndur) {
    HString& s = st->s[si];
    const int N = s.n_nodes;
    const float pos = fmaxf(1e-3f, fminf(1.0f - 1e-3f, pos_n));
    const float amp = fmaxf(-0.006f, fminf(0.006f, amp_in));
    int ip = (int)roundf(pos * (float)(N - 1));
    if (ip <= 0) ip = 1;
    if (ip >= N - 1) ip = N - 2;
    const float ll = fmaxf((float)ip * s.ds, s.ds);
    const float rl = fmaxf((float)(N - 1 - ip) * s.ds, s.ds);
    float F = s.T0 * fabsf(amp) * (1.0f / ll + 1.0f / rl);
    if (amp < 0.0f) F = -F;
    int nd = (ndur > 0) ? ndur : (int)roundf(0.020f * st->sr);
    if (nd < 4) nd = 4;
    const float sc = s.dt_s * s.dt_s / (s.mu * s.ds);
    s.fbuf.resize(nd);
    for (int q = 0; q < nd; ++q) {
        const float t = (float)q / (float)(nd - 1);
        const float env = 0.5f * (1.0f - cosf(kPi * t));
        s.fbuf[q] = F * env * sc;
    }
    s.fnode = ip;
    s.fonset = st->cur_sample;
    return CE_OK;
}

void fire_plucks(AcousticCoEvolverState* st, int sample_idx) {   // :2237-2244
    while (st->cursor < (int)st->ev.size() && st->ev[st->cursor].onset <= sample_idx) {
        const HPluck e = st->ev[st->cursor++];
        pluck_host(st, e.si, e.pos, e.amp, e.ndur);
    }
}

// Per-sample string force of one sample (:1919-1931).
void plan_sample(AcousticCoEvolverState* st, BlockPlan& bp, int q, int sample_abs) {
    for (int si = 0; si < st->ns; ++si) {
        const HString& s = st->s[si];
        const int bi = sample_abs - s.fonset;
        const float ft = (bi >= 0 && bi < (int)s.fbuf.size()) ? s.fbuf[bi] : 0.0f;
        const int tot = st->K * s.n_sub;
        bp.sub[(size_t)q * st->ns + si] = (tot > 0 && ft != 0.0f) ? ft / (float)tot : 0.0f;
        bp.node[(size_t)q * st->ns + si] = s.fnode;
    }
}

struct HostSnap {
    int cursor, cur_sample;
    std::vector<std::vector<float>> fbuf;
    std::vector<int> fnode, fonset;
};

HostSnap snap(const AcousticCoEvolverState* st) {
    HostSnap h;
    h.cursor = st->cursor;
    h.cur_sample = st->cur_sample;
    for (const auto& s : st->s) {
        h.fbuf.push_back(s.fbuf);
        h.fnode.push_back(s.fnode);
        h.fonset.push_back(s.fonset);
    }
    return h;
}

void restore(AcousticCoEvolverState* st, const HostSnap& h) {
    st->cursor = h.cursor;
    st->cur_sample = h.cur_sample;
    for (int i = 0; i < st->ns; ++i) {
        st->s[i].fbuf = h.fbuf[i];
        st->s[i].fnode = h.fnode[i];
        st->s[i].fonset = h.fonset[i];
    }
}

enum Mode { kStep = 0, kWorker = 1, kDrive = 2 };

// Host side of n samples: plucks fired, per-sample forces.  mode kStep uses
// the absolute sample index (coevolver_step), kWorker the job-relative index
// (the async worker, :2338), kDrive none (block-with-drive passes -1).
void plan_block(AcousticCoEvolverState* st, BlockPlan& bp, int n, Mode mode, int job_off) {
    bp.sub.assign((size_t)n * st->ns, 0.0f);
    bp.node.assign((size_t)n * st->ns, 0);
    const int base = st->written;
    for (int q = 0; q < n; ++q) {
        const int sample_idx = base + q;
        if (mode == kStep) st->cur_sample = sample_idx;
        fire_plucks(st, sample_idx);
        const int sabs = mode == kStep ? sample_idx : (mode == kWorker ? job_off + q : -1);
        plan_sample(st, bp, q, sabs);
    }
}

// Every buffer k_coevo dereferences (blk_dval only in the block-drive form).
bool co_pointers_set(const CoConst& k) {
    const void* p[] = {k.str, k.sw, k.u, k.neck, k.nidx, k.nwgt, k.ngrp, k.sbk, k.sad_idx,
                       k.sad_wgt, k.basis, k.modal, k.K0, k.K1, k.magmap, k.pk_type, k.pk_mask,
                       k.pk_sens, k.mPi, k.mPw, k.mVi, k.mVw, k.mic_par, k.pk_out, k.mic_out,
                       k.blk_sub, k.blk_node, k.blk_dnode, k.status};
    for (const void* q : p)
        if (!q) return false;
    return k.nb == 0 || k.bgrp != nullptr;
}

hipStream_t co_stream(const AcousticCoEvolverState* st) {
    if (st->fdtd) return st->fdtd->stream;
    AmrCoView v;
    return (st->amr && amr_co_view(st->amr, &v) == SK_OK) ? v.stream : nullptr;
}

int ensure_blk(AcousticCoEvolverState* st, int n) {
    const size_t need = (size_t)n * std::max(st->ns, 1);
    if ((int)need <= st->blk_cap) return CE_OK;
    if (st->d_blk_sub) (void)hipFree(st->d_blk_sub);
    if (st->d_blk_node) (void)hipFree(st->d_blk_node);
    if (st->d_blk_dval) (void)hipFree(st->d_blk_dval);
    st->d_blk_sub = nullptr; st->d_blk_node = nullptr; st->d_blk_dval = nullptr;
    st->blk_cap = 0;
    if (dalloc(&st->d_blk_sub, need) != hipSuccess || dalloc(&st->d_blk_node, need) != hipSuccess ||
        dalloc(&st->d_blk_dval, need) != hipSuccess) {
        afdtd_set_error("coevolver: block buffers");
        return CE_ERR_FDTD_STEP;
    }
    st->blk_cap = (int)need;
    return CE_OK;
}

// Run n (<= kChunk) planned samples on the device.  Returns CE_OK or an
// error; *done = samples completed.
int run_block(AcousticCoEvolverState* st, const BlockPlan& bp, int n, bool drive, int* done,
              float* ms_total) {
    *done = 0;
    if (n <= 0) return CE_OK;
    int rc = ensure_blk(st, n);
    if (rc) return rc;
    AcousticFDTDState* f = st->fdtd;
    const hipStream_t sm = co_stream(st);
    const size_t cnt = (size_t)n * st->ns;
    if (cnt) {
        if (hipMemcpyAsync(st->d_blk_sub, bp.sub.data(), cnt * sizeof(float),
                           hipMemcpyHostToDevice, sm) != hipSuccess ||
            hipMemcpyAsync(st->d_blk_node, bp.node.data(), cnt * sizeof(int),
                           hipMemcpyHostToDevice, sm) != hipSuccess)
            return CE_ERR_FDTD_STEP;
        if (drive &&
            (hipMemcpyAsync(st->d_blk_dval, bp.dval.data(), cnt * sizeof(float),
                            hipMemcpyHostToDevice, sm) != hipSuccess ||
             hipMemcpyAsync(st->d_blk_dnode, bp.dnode.data(), st->ns * sizeof(int),
                            hipMemcpyHostToDevice, sm) != hipSuccess))
            return CE_ERR_FDTD_STEP;
    }
    // damping scale of this block (:1924)
    if (st->gamma_dev != st->gamma_scale) {
        std::vector<CoStr> hs(st->ns);
        if (st->ns && hipMemcpy(hs.data(), st->d_str, st->ns * sizeof(CoStr),
                                hipMemcpyDeviceToHost) != hipSuccess)
            return CE_ERR_FDTD_STEP;
        for (int i = 0; i < st->ns; ++i) hs[i].gamma_eff = st->s[i].gamma * st->gamma_scale;
        if (st->ns && hipMemcpy(st->d_str, hs.data(), st->ns * sizeof(CoStr),
                                hipMemcpyHostToDevice) != hipSuccess)
            return CE_ERR_FDTD_STEP;
        st->gamma_dev = st->gamma_scale;
    }
    const int K = st->K, nsteps = n * K;
    // modal schedule: sub-step q projects when the running counter reaches the stride
    std::vector<uint8_t> modal(nsteps, 0);
    int mc = st->modal_counter;
    for (int q = 0; q < nsteps; ++q) {
        if (++mc >= st->modal_stride) { mc = 0; modal[q] = 1; }
    }
    CoConst kc = st->k;
    kc.blk_sub = st->d_blk_sub;
    kc.blk_node = st->d_blk_node;
    kc.blk_dval = drive ? st->d_blk_dval : nullptr;
    if (!co_pointers_set(kc)) {
        afdtd_set_error("coevolver: a device buffer of the step kernel is not allocated");
        return CE_ERR_FDTD_STEP;
    }
    const int out0 = st->out_pos;
    const size_t lds = st->lds_bytes;
    // uniform body: the sweep of each step forms the next sub-step's neck sum
    const bool side = !st->amr && kc.neck_on && st->d_npre && f->fused_geo &&
                      (f->family == FDTD_KERNEL_FUSED || f->family == FDTD_KERNEL_FUSED2);
    auto launch = [&](int q, int step, const float* w, float* ext, const float* P,
                      const float* const* V) -> bool {
        CoLaunch a;
        a.npre = side && q > 0;
        a.sample = q / K;
        a.sub = q % K;
        a.step = step;
        a.post = q > 0;
        a.modal = q > 0 ? modal[q - 1] : 0;
        a.sample_end = q > 0 && ((q - 1) % K) == K - 1;
        a.out_pos = q > 0 ? (out0 + (q - 1) / K) % kRing : 0;
        a.run = q < nsteps;
        a.w = w;
        a.ext = ext;
        a.P = P;
        for (int ax = 0; ax < 3; ++ax) a.V[ax] = V[ax];
        if (!a.post && !a.run) return true;
        hipLaunchKernelGGL(k_coevo, dim3(1), dim3(kThreads), lds, sm, kc, a);
        return hipGetLastError() == hipSuccess;
    };
    int done_steps = nsteps;
    bool ok = true;
    if (st->amr) {
        // AMR body: the hook sees relative step indices; a failed hook or a
        // diverged step ends the call with the steps before it completed
        AmrStepHook hook = [&](const AmrHookArgs& h) -> int {
            const float* V[3] = {h.v, nullptr, nullptr};
            return launch(h.q, h.q, h.w, h.ext, h.p, V) ? SK_OK : AMR_ERR_DEVICE;
        };
        int hook_failed = 0;
        const int arc = amr_step_hooked(st->amr, nsteps, hook, ms_total, &done_steps, &hook_failed);
        if (arc != SK_OK && arc != SK_ERR_DIVERGED) return CE_ERR_FDTD_STEP;
        ok = arc == SK_OK;
    } else {
        StepHook hook = [&](int q, int s) -> int {
            const float* V[3] = {(const float*)f->V[0], (const float*)f->V[1],
                                 (const float*)f->V[2]};
            if (!launch(q, s, f->d_w[0], f->d_ext, (const float*)f->P[f->cur], V))
                return FDTD_ERR_NULL;
            return FDTD_OK;
        };
        int status[2] = {0, 0};
        const int start_step = f->step_count;
        if (side) {
            f->side_idx = st->d_nidx;
            f->side_wgt = st->d_nwgt;
            f->side_n = st->nn;
            f->side_out = st->d_npre;
        }
        const int frc = afdtd_step_hooked(f, nsteps, hook, ms_total, status);
        f->side_n = 0;
        f->side_out = nullptr;
        if (frc == FDTD_ERR_UNSTABLE) done_steps = status[1] - start_step;
        else if (frc != FDTD_OK) return CE_ERR_FDTD_STEP;
        ok = frc == FDTD_OK;
    }
    // sub-steps that finished (saddle gathered, modal counter advanced)
    for (int q = 0; q < done_steps; ++q)
        if (++st->modal_counter >= st->modal_stride) st->modal_counter = 0;
    const int samples = ok ? n : done_steps / K;
    st->out_pos = (out0 + samples) % kRing;
    st->written += samples;
    *done = samples;
    return ok ? CE_OK : CE_ERR_UNSTABLE;
}

// n samples in chunks; mode as plan_block.  mic0 (block-drive form) gets
// mic 0 of every completed sample.
int step_samples(AcousticCoEvolverState* st, int n, Mode mode, const float* const* blocks,
                 float pos_n, float* mic0, int job_off, float* ms_total, int* done_out) {
    int done_all = 0;
    float ms_sum = 0.0f;
    BlockPlan bp;
    if (mode == kDrive) {                                        // :2272-2283
        bp.dnode.assign(st->ns, 1);
        for (int si = 0; si < st->ns; ++si) {
            const HString& s = st->s[si];
            const int N = s.n_nodes;
            const float pos = fmaxf(0.0f, fminf(1.0f, pos_n));
            int i0 = (int)roundf(pos * (float)(N - 1));
            if (i0 <= 0) i0 = 1;
            if (i0 >= N - 1) i0 = N - 2;
            bp.dnode[si] = i0;
        }
    }
    int rc = CE_OK;
    while (done_all < n) {
        const int m = std::min(kChunk, n - done_all);
        const HostSnap h = snap(st);
        const int base = st->written;
        plan_block(st, bp, m, mode, job_off + done_all);
        if (mode == kDrive) {
            bp.dval.assign((size_t)m * st->ns, 0.0f);
            if (blocks)
                for (int si = 0; si < st->ns; ++si) {
                    if (!blocks[si]) continue;
                    const HString& s = st->s[si];
                    const float sc = s.dt_s * s.dt_s / (s.mu * s.ds);
                    for (int q = 0; q < m; ++q) {
                        const float fv = blocks[si][done_all + q];
                        bp.dval[(size_t)q * st->ns + si] = fv != 0.0f ? fv * sc : 0.0f;
                    }
                }
        }
        const int ring0 = st->out_pos;
        int done = 0;
        float ms = 0.0f;
        rc = run_block(st, bp, m, mode == kDrive, &done, ms_total ? &ms : nullptr);
        ms_sum += ms;
        if (mic0 && done > 0) {
            std::vector<float> ring(kRing);
            if (hipMemcpy(ring.data(), st->d_mic_out, kRing * sizeof(float),
                          hipMemcpyDeviceToHost) != hipSuccess)
                return CE_ERR_FDTD_MIC_P;
            for (int q = 0; q < done; ++q) mic0[done_all + q] = ring[(ring0 + q) % kRing];
        }
        if (rc != CE_OK && done < m) {
            // the host side stops where the device stopped: replay the
            // plucks of the completed samples and of the failing one
            restore(st, h);
            const int keep = st->written;
            st->written = base;
            BlockPlan tmp;
            plan_block(st, tmp, std::min(done + 1, m), mode, job_off + done_all);
            st->written = keep;
        }
        done_all += done;
        if (rc != CE_OK) break;
    }
    if (ms_total) *ms_total = ms_sum;
    if (done_out) *done_out = done_all;
    return rc;
}

int copy_ring(AcousticCoEvolverState* st, const float* ring, float* out, int n) {   // :2511-2518
    if (n > kRing) return CE_ERR_DIM;
    if (n <= 0) return CE_OK;
    std::vector<float> h(kRing);
    if (hipMemcpy(h.data(), ring, kRing * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return CE_ERR_NULL;
    const int start = (st->out_pos - n + kRing) % kRing;
    for (int i = 0; i < n; ++i) out[i] = h[(start + i) % kRing];
    return CE_OK;
}

int read_strings(const AcousticCoEvolverState* st, std::vector<float>& u) {
    u.resize((size_t)4 * st->tot_nodes);
    if (st->tot_nodes == 0) return CE_OK;
    if (hipMemcpy(u.data(), st->d_u, u.size() * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return CE_ERR_NULL;
    return CE_OK;
}

}  // namespace

namespace {

// Set-up shared by both bodies (acoustic_coevolver.cpp:1063-1381 and
// :1480-1745): strings, neck constants, pickups and their GEMV rows, mics,
// saddle taps, per-string bridge kernels, modal basis, force groups, device
// state and the kernel's constants.  The body, the grid fields (Nx, Ny, Nz,
// dx, origin; the plate grid with Nz = 1 on an AMR body), dt_fdtd, K and the
// bridge (and, on an AMR body, neck) plate lists are in place.  Returns
// nullptr, or the failing stage.
const char* co_setup(AcousticCoEvolverState* st, int n_strings, const CoEvolverStringDef* sd,
                     int n_pickups, const CoEvolverPickupDef* pd, int n_mics,
                     const CoEvolverMicDef* md, const CoEvolverBodyDef* body, float bridge_sigma) {
    AcousticFDTDState* f = st->fdtd;
    const int nplate = f ? f->Nplate : st->Nx * st->Ny;
    st->ns = n_strings;
    st->s.resize(n_strings);
    for (int i = 0; i < n_strings; ++i)
        if (!string_init(st->s[i], sd[i], st->dt_fdtd)) return "string_create";
    for (int i = 0; i < n_strings; ++i) {                                  // :1079-1101
        if (sd[i].neck_freq_hz > 0.0f && sd[i].neck_mass_kg > 0.0f) {
            st->n_m = sd[i].neck_mass_kg;
            const float om = 2.0f * kPi * sd[i].neck_freq_hz;
            st->n_k = st->n_m * om * om;
            const float Q = (sd[i].neck_Q > 0.0f) ? sd[i].neck_Q : 30.0f;
            st->n_c = st->n_m * om / Q;
            const float oj = 2.0f * kPi * 95.0f;
            st->nb_k = st->n_m * oj * oj;
            st->nb_c = st->n_m * oj / 2.0f;
            break;
        }
    }
    if (st->amr) {
        if (st->n_m > 0.0f && st->nn == 0) st->n_m = 0.0f;          // :1510-1512
    } else if (st->n_m > 0.0f && neck_cells(st, body->plate_active, 0.030f) != CE_OK) {
        return "neck_cells";
    }
    // node / segment offsets
    int no = 0, so = 0;
    for (auto& s : st->s) {
        s.node_off = no;
        s.seg_off = so;
        no += s.n_nodes;
        so += s.n_segs;
    }
    st->tot_nodes = no;
    st->tot_seg = so;
    if (no > kMaxNodes) {
        afdtd_set_error("coevolver_create: more than 4096 string nodes in total");
        return "string_nodes";
    }
    // pickups (:1107-1132, GEMV rows :1248-1306)
    st->npk = n_pickups;
    std::vector<float> K0, K1;
    std::vector<int> magmap;
    for (int p = 0; p < n_pickups; ++p) {
        st->pk_type.push_back(pd[p].type);
        st->pk_mask.push_back(pd[p].string_mask);
        st->pk_sens.push_back(pd[p].sensitivity);
    }
    for (int p = 0; p < n_pickups; ++p)
        if (pd[p].type != PICKUP_PIEZO) ++st->nmag;
    if (st->nmag > 0 && st->tot_seg > 0) {
        K0.assign((size_t)st->nmag * st->tot_seg, 0.0f);
        K1.assign((size_t)st->nmag * st->tot_seg, 0.0f);
        int r = 0;
        for (int p = 0; p < n_pickups; ++p) {
            if (pd[p].type == PICKUP_PIEZO) continue;
            magmap.push_back(p);
            for (int i = 0; i < n_strings; ++i) {
                if (!(pd[p].string_mask & (1u << (unsigned)i))) continue;
                std::vector<float> Bn, Bb;
                bkernels(pd[p], st->s[i], Bn, Bb);
                for (int q = 0; q < st->s[i].n_segs; ++q) {
                    K0[(size_t)r * st->tot_seg + st->s[i].seg_off + q] = Bn[q] * st->s[i].ds;
                    K1[(size_t)r * st->tot_seg + st->s[i].seg_off + q] = Bb[q] * st->s[i].ds;
                }
            }
            ++r;
        }
    } else {
        st->nmag = std::min(st->nmag, 0);
    }
    // mics (:1135-1246)
    st->nmic = n_mics;
    std::vector<float> mic_par((size_t)6 * n_mics);
    st->mic_g.resize((size_t)3 * n_mics);
    for (int m = 0; m < n_mics; ++m) {
        for (int q = 0; q < 3; ++q)
            st->mic_g[3 * m + q] = (md[m].pos[q] - st->origin[q]) / st->dx;
        mic_par[6 * m + 0] = md[m].polar_a;
        mic_par[6 * m + 1] = md[m].polar_b;
        mic_par[6 * m + 2] = md[m].gain;
        for (int q = 0; q < 3; ++q) mic_par[6 * m + 3 + q] = md[m].axis[q];
        if (md[m].polar_b != 0.0f) st->any_vel = 1;
    }
    std::vector<int> mPi((size_t)8 * n_mics), mVi;
    std::vector<float> mPw((size_t)8 * n_mics), mVw;
    if (n_mics > 0 && st->amr) {                                          // :1613-1646
        std::vector<float> pos((size_t)3 * n_mics), axis((size_t)3 * n_mics);
        for (int m = 0; m < n_mics; ++m)
            for (int q = 0; q < 3; ++q) {
                pos[3 * m + q] = md[m].pos[q];
                axis[3 * m + q] = md[m].axis[q];
            }
        if (amr_setup_mic_samplers(st->amr, n_mics, pos.data(), axis.data(), st->any_vel) != SK_OK)
            return "setup_mic_samplers";
    } else if (n_mics > 0) {
        if (fdtd_precompute_pressure_samplers(f, n_mics, st->mic_g.data(), mPi.data(),
                                              mPw.data()) != FDTD_OK)
            return "setup_mic_samplers";
        mVi.assign((size_t)24 * n_mics, 0);
        mVw.assign((size_t)24 * n_mics, 0.0f);
        if (st->any_vel) {
            const size_t o = (size_t)8 * n_mics;
            if (fdtd_precompute_velocity_samplers(f, n_mics, st->mic_g.data(), mVi.data(),
                                                  mVw.data(), mVi.data() + o, mVw.data() + o,
                                                  mVi.data() + 2 * o,
                                                  mVw.data() + 2 * o) != FDTD_OK)
                return "setup_mic_samplers";
            for (int ax = 0; ax < 3; ++ax)
                for (size_t q = 0; q < o; ++q)
                    mVi[ax * o + q] = packed_to_padded_v(ax, mVi[ax * o + q], st->Nx, st->Ny,
                                                         st->Nz);
        }
    }
    // saddle bilinear taps on the plate (:1160-1183)
    std::vector<int> sad_idx((size_t)4 * n_strings);
    std::vector<float> sad_wgt((size_t)4 * n_strings);
    for (int i = 0; i < n_strings; ++i) {
        const HString& s = st->s[i];
        const int Nn = s.n_nodes;
        float gi = (s.xyz[(Nn - 1) * 3 + 0] - st->origin[0]) / st->dx;
        float gj = (s.xyz[(Nn - 1) * 3 + 1] - st->origin[1]) / st->dx;
        if (gi < 0.0f) gi = 0.0f;
        if (gi > (float)(st->Nx - 1)) gi = (float)(st->Nx - 1);
        if (gj < 0.0f) gj = 0.0f;
        if (gj > (float)(st->Ny - 1)) gj = (float)(st->Ny - 1);
        const int i0 = (int)gi, i1 = (i0 < st->Nx - 1) ? i0 + 1 : i0;
        const int j0 = (int)gj, j1 = (j0 < st->Ny - 1) ? j0 + 1 : j0;
        const float fi = gi - (float)i0, fj = gj - (float)j0;
        sad_idx[i * 4 + 0] = j0 + st->Ny * i0;
        sad_idx[i * 4 + 1] = j0 + st->Ny * i1;
        sad_idx[i * 4 + 2] = j1 + st->Ny * i0;
        sad_idx[i * 4 + 3] = j1 + st->Ny * i1;
        sad_wgt[i * 4 + 0] = (1.0f - fi) * (1.0f - fj);
        sad_wgt[i * 4 + 1] = fi * (1.0f - fj);
        sad_wgt[i * 4 + 2] = (1.0f - fi) * fj;
        sad_wgt[i * 4 + 3] = fi * fj;
    }
    // per-string bridge kernels (:1322-1375)
    std::vector<float> sbk;
    if (n_strings > 0 && st->nb > 0) {
        const int nb = st->nb;
        sbk.assign((size_t)n_strings * nb, 0.0f);
        const float sigma = bridge_sigma;
        const float i2s2 = 0.5f / (sigma * sigma);
        for (int i = 0; i < n_strings; ++i) {
            const HString& s = st->s[i];
            const int Nn = s.n_nodes;
            const float sx = s.xyz[(Nn - 1) * 3 + 0], sy = s.xyz[(Nn - 1) * 3 + 1];
            float* row = &sbk[(size_t)i * nb];
            float ws = 0.0f;
            for (int cix = 0; cix < nb; ++cix) {
                const int flat = st->bidx[cix];
                const int t = flat / st->Nz;
                const int j = t % st->Ny, ii = t / st->Ny;
                const float cx = (float)ii * st->dx + st->origin[0];
                const float cy = (float)j * st->dx + st->origin[1];
                const float d2x = sx - cx, d2y = sy - cy;
                row[cix] = expf(-(d2x * d2x + d2y * d2y) * i2s2);
                ws += row[cix];
            }
            if (ws > 1e-12f)
                for (int cix = 0; cix < nb; ++cix) row[cix] /= ws;
        }
    }
    // modal basis, node-major for coalesced mode reads (:494-505)
    std::vector<float> basis((size_t)st->tot_nodes * kModes);
    for (const auto& s : st->s) {
        const int Nn = s.n_nodes;
        const float inv = 1.0f / (float)(Nn > 1 ? Nn - 1 : 1);
        for (int k = 0; k < kModes; ++k) {
            const float kpi = (float)(k + 1) * kPi;
            for (int i = 0; i < Nn; ++i)
                basis[(size_t)(s.node_off + i) * kModes + k] = sinf(kpi * (float)i * inv);
        }
    }
    // neck force groups over the carved plate mask the body FDTD holds
    std::vector<int> ngrp;
    int nngrp = 0;
    if (st->nn > 0)
        nngrp = group_entries(st->nidx.data(), st->nn, 1, st->amr ? nullptr : f->h_pa.data(),
                              nplate, ngrp);
    // AMR body: the bridge plate entries grouped by node (the uniform body's
    // groups live in its FDTD state)
    std::vector<int> bgrp;
    int nbgrp = f ? f->nbgrp : 0;
    if (st->amr && st->nb > 0)
        nbgrp = group_entries(st->bidx.data(), st->nb, 1, nullptr, nplate, bgrp);
    // device state
    std::vector<CoStr> hs(n_strings);
    std::vector<float> sw((size_t)kSw * std::max(n_strings, 1), 0.0f);
    for (int i = 0; i < n_strings; ++i) {
        const HString& s = st->s[i];
        CoStr& d = hs[i];
        d.n_nodes = s.n_nodes; d.node_off = s.node_off; d.seg_off = s.seg_off; d.n_sub = s.n_sub;
        d.ds = s.ds; d.L = s.L; d.T0 = s.T0; d.EA = s.EA; d.mu = s.mu; d.dt_s = s.dt_s;
        d.biharm = s.biharm; d.gamma_eff = s.gamma;
        d.b_m = s.b_m; d.b_k = s.b_k; d.b_c = s.b_c; d.inv_dt_s = 1.0f / s.dt_s;
        sw[(size_t)i * kSw + kTeff] = s.T0;
    }
    st->gamma_dev = 1.0f;
    std::vector<float> zeros_u((size_t)4 * std::max(st->tot_nodes, 1), 0.0f);
    std::vector<float> neck(4, 0.0f);
#define UP(p, v) if (dupload(&st->p, v) != hipSuccess) return "device_alloc"
    UP(d_str, hs);
    UP(d_sw, sw);
    UP(d_u, zeros_u);
    UP(d_neck, neck);
    UP(d_nidx, st->nidx);
    UP(d_nwgt, st->nwgt);
    UP(d_ngrp, ngrp);
    UP(d_sbk, sbk);
    UP(d_sad_idx, sad_idx);
    UP(d_sad_wgt, sad_wgt);
    UP(d_basis, basis);
    UP(d_K0, K0);
    UP(d_K1, K1);
    UP(d_magmap, magmap);
    UP(d_pk_type, st->pk_type);
    UP(d_pk_mask, st->pk_mask);
    UP(d_pk_sens, st->pk_sens);
    UP(d_mPi, mPi);
    UP(d_mPw, mPw);
    UP(d_mVi, mVi);
    UP(d_mVw, mVw);
    UP(d_mic_par, mic_par);
    UP(d_bwgt, st->bwgt);
    UP(d_bgrp, bgrp);
    if (!st->amr) {                       // neck sum slot: +0 and a tag no step has
        std::vector<float> np0(2, 0.0f);
        const int none = INT_MIN;
        memcpy(&np0[1], &none, sizeof(int));
        UP(d_npre, np0);
    }
#undef UP
    {
        std::vector<float> modal0((size_t)2 * kModes * std::max(n_strings, 1), 0.0f);
        std::vector<float> pk0((size_t)kRing * std::max(n_pickups, 1), 0.0f);
        std::vector<float> mic0((size_t)kRing * std::max(n_mics, 1), 0.0f);
        std::vector<int> dn(std::max(n_strings, 1), 1);
        if (dupload(&st->d_modal, modal0) != hipSuccess || dupload(&st->d_pk_out, pk0) != hipSuccess ||
            dupload(&st->d_mic_out, mic0) != hipSuccess || dupload(&st->d_blk_dnode, dn) != hipSuccess)
            return "device_alloc";
    }
    // LDS: 8 words per string node, then neck products, fpl[64], drive[nb], flags
    // (the post phase's 2T modal staging and 4S pickup rows fit in the 8T string words)
    // (an AMR body's mic stencil sums take 4 * kThreads + 4 * n_mics words of it)
    size_t lds_str = (size_t)8 * st->tot_nodes;
    if (st->amr && n_mics > 0) lds_str = std::max(lds_str, (size_t)4 * kThreads + 4 * (size_t)n_mics);
    st->lds_bytes = sizeof(float) * (lds_str + st->nn + kMaxStrings + st->nb + 8 +
                                     (3 + kSw + 2) * (size_t)std::max(n_strings, 1));
    // the per-string bridge kernels in LDS when they fit beside the rest
    const size_t nsbk = (size_t)n_strings * st->nb;
    const bool sbk_lds = nsbk > 0 && st->lds_bytes + sizeof(float) * nsbk + 16384 <= (size_t)kLdsLimit;
    if (sbk_lds) st->lds_bytes += sizeof(float) * nsbk;
    // deferred output sums: two wavefronts without a string (n_strings <= 13),
    // one lane per sum (at most 32 magnetic pickups, 16 AMR mics), and LDS
    // for the products
    AmrCoView av{};
    if (st->amr && amr_co_view(st->amr, &av) != SK_OK) return "device_alloc";
    size_t ndef = (size_t)2 * st->tot_seg * st->nmag;
    if (st->amr && n_mics > 0) ndef += (size_t)av.n_mp + 3 * (size_t)av.n_mv + 4 * (size_t)n_mics;
    const bool defer = ndef > 0 && n_strings <= kThreads / 64 - 3 && st->nmag <= 32 &&
                       !(st->amr && n_mics > 16) &&
                       st->lds_bytes + sizeof(float) * ndef <= (size_t)kLdsLimit;
    if (defer) st->lds_bytes += sizeof(float) * ndef;
    if (st->lds_bytes > (size_t)kLdsLimit) {
        afdtd_set_error("coevolver_create: strings, neck and bridge kernels exceed the LDS of one "
                        "workgroup");
        return "lds_limit";
    }
    if (hipFuncSetAttribute((const void*)k_coevo, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)st->lds_bytes) != hipSuccess)
        return "device_alloc";
    CoConst& k = st->k;
    k.ns = n_strings; k.npk = n_pickups; k.nmic = n_mics; k.nmag = st->nmag;
    k.tot_seg = st->tot_seg; k.tot_nodes = st->tot_nodes; k.nb = st->nb; k.nbgrp = nbgrp;
    k.nn = st->nn; k.nngrp = nngrp; k.neck_on = (st->n_m > 0.0f && st->nn > 0) ? 1 : 0;
    k.any_vel = st->any_vel; k.Nplate = nplate; k.clear_ext = st->amr ? 0 : 1;
    k.amr = st->amr ? 1 : 0;
    k.defer = defer ? 1 : 0;
    k.sad_early = n_strings <= kThreads / 64 - 4 ? 1 : 0;
    k.sbk_lds = sbk_lds ? 1 : 0;
    k.anp = av.n_mp;
    k.anv = av.n_mv;
    k.st_ok = st->amr ? INT_MAX : 0;
    k.lds_str = (int)lds_str;
    float dt_b = st->dt_fdtd;
    if (dt_b <= 0.0f) dt_b = st->dt_audio;
    k.dt_b = dt_b;
    k.inv_dx2 = f ? 1.0f / (f->dx * f->dx) : 0.0f;
    k.rhoc = st->rho_air * st->c;
    k.n_m = st->n_m; k.n_c = st->n_c + st->nb_c; k.n_k = st->n_k + st->nb_k;
    k.nb_k = st->nb_k; k.nb_c = st->nb_c;
    k.str = st->d_str; k.sw = st->d_sw; k.u = st->d_u; k.neck = st->d_neck;
    k.nidx = st->d_nidx; k.nwgt = st->d_nwgt; k.ngrp = st->d_ngrp; k.sbk = st->d_sbk;
    k.bgrp = st->amr ? st->d_bgrp : f->d_bgrp; k.sad_idx = st->d_sad_idx; k.sad_wgt = st->d_sad_wgt;
    k.basis = st->d_basis; k.modal = st->d_modal; k.K0 = st->d_K0; k.K1 = st->d_K1;
    k.magmap = st->d_magmap; k.pk_type = st->d_pk_type; k.pk_mask = st->d_pk_mask;
    k.pk_sens = st->d_pk_sens; k.mPi = st->d_mPi; k.mPw = st->d_mPw; k.mVi = st->d_mVi;
    k.mVw = st->d_mVw; k.mic_par = st->d_mic_par; k.pk_out = st->d_pk_out;
    k.mic_out = st->d_mic_out; k.blk_dnode = st->d_blk_dnode;
    k.bwgt = st->d_bwgt;
    k.npre = st->d_npre;
    if (st->amr) {
        const AmrCoView& v = av;
        k.status = v.status;
        k.amp_start = v.mp_start; k.amp_cell = v.mp_cell; k.amp_w = v.mp_w;
        k.amv_start = v.mv_start; k.amv_face = v.mv_face;
        for (int ax = 0; ax < 3; ++ax) k.amv_w[ax] = v.mv_w[ax];
    } else {
        k.status = f->d_status;
        f->hook_clears_ext = 1;
    }
    if (ensure_blk(st, 64) != CE_OK) return "device_alloc";
    return nullptr;
}

}  // namespace

extern "C" {

SK_API AcousticCoEvolverState* coevolver_create(int n_strings, const CoEvolverStringDef* sd,
                                                int n_pickups, const CoEvolverPickupDef* pd,
                                                int n_mics, const CoEvolverMicDef* md,
                                                const CoEvolverBodyDef* body, float sample_rate,
                                                int modal_stride, float force_scale) {
    (void)force_scale;                                                   // :1043-1044
    if (!sd || !pd || !md || !body) return nullptr;                       // :999-1002
    if (n_strings < 0 || n_pickups < 0 || n_mics < 0) return nullptr;
    if (sample_rate <= 0 || modal_stride <= 0) return nullptr;
    if (n_strings > kMaxStrings) {
        afdtd_set_error("coevolver_create: at most 64 strings");
        return nullptr;
    }
    if (!body->cell_type || !body->plate_active || body->Nx <= 0 || body->Ny <= 0 ||
        body->Nz <= 0 || body->dx <= 0.0f)
        return nullptr;
    AcousticCoEvolverState* st = new (std::nothrow) AcousticCoEvolverState();
    if (!st) return nullptr;
    auto fail = [&]() -> AcousticCoEvolverState* {
        co_free(st);
        delete st;
        return nullptr;
    };
    {
        std::vector<uint8_t> pa = carve_soundhole(body);
        st->fdtd = fdtd_create(body->Nx, body->Ny, body->Nz, body->dx, body->c, body->rho_air,
                               body->cell_type, body->plate_iz, pa.data(),
                               body->plate_mass_density, body->plate_stiffness_D,
                               body->plate_alpha_M, body->plate_beta_K, body->n_pml);
    }
    if (!st->fdtd) return fail();
    AcousticFDTDState* f = st->fdtd;
    st->Nx = body->Nx; st->Ny = body->Ny; st->Nz = body->Nz; st->dx = body->dx;
    st->rho_air = body->rho_air; st->c = body->c; st->plate_iz = body->plate_iz;
    for (int q = 0; q < 3; ++q) st->origin[q] = body->origin[q];
    st->sr = sample_rate;
    st->dt_audio = 1.0f / sample_rate;
    st->dt_fdtd = fdtd_get_dt(f);
    st->modal_stride = modal_stride;
    st->K = (int)ceilf(st->dt_audio / st->dt_fdtd);
    if (st->K < 1) st->K = 1;
    if (st->K > kMaxSub) st->K = kMaxSub;
    if (body->n_bridge_src > 0 && body->bridge_src_xyz)
        bridge_cells(st, body->n_bridge_src, body->bridge_src_xyz, body->bridge_sigma);
    if (st->nb > 0 &&
        fdtd_set_bridge_sources(f, st->nb, st->bidx.data(), st->bwgt.data()) != FDTD_OK)
        return fail();
    const float sigma = (body->bridge_sigma > 0.0f) ? body->bridge_sigma : st->dx;   // :1322-1375
    if (co_setup(st, n_strings, sd, n_pickups, pd, n_mics, md, body, sigma)) return fail();
    return st;
}

SK_API AcousticCoEvolverState* coevolver_create_amr(int n_strings,
                                                    const CoEvolverStringDef* sd, int n_pickups,
                                                    const CoEvolverPickupDef* pd, int n_mics,
                                                    const CoEvolverMicDef* md,
                                                    const AMRCoevolverDescriptor* d,
                                                    float sample_rate, int modal_stride) {
    g_amr_err = "";                                                      // :1398-1413
    if (n_strings < 0 || n_pickups < 0 || n_mics < 0) {
        g_amr_err = "invalid_counts";
        return nullptr;
    }
    if ((n_strings > 0 && !sd) || (n_pickups > 0 && !pd) || (n_mics > 0 && !md) || !d) {
        g_amr_err = "invalid_null_args";
        return nullptr;
    }
    if (sample_rate <= 0 || modal_stride <= 0) {
        g_amr_err = "invalid_timing_args";
        return nullptr;
    }
    if (n_strings > kMaxStrings) {
        afdtd_set_error("coevolver_create_amr: at most 64 strings");
        g_amr_err = "strings_alloc";
        return nullptr;
    }
    AcousticCoEvolverState* st = new (std::nothrow) AcousticCoEvolverState();
    if (!st) return nullptr;
    auto fail = [&](const char* stage) -> AcousticCoEvolverState* {
        g_amr_err = stage;
        co_free(st);
        delete st;
        return nullptr;
    };
    st->amr = amr_state_from_descriptor(d);                             // :1422-1424
    if (!st->amr) return fail("pressure_backend_create");
    st->Nx = d->plate_Nx; st->Ny = d->plate_Ny; st->Nz = 1; st->dx = d->plate_dx;   // :1429-1437
    for (int q = 0; q < 3; ++q) st->origin[q] = d->plate_origin[q];
    st->rho_air = (float)d->rho_air;
    st->c = (float)d->c;
    st->sr = sample_rate;
    st->dt_audio = 1.0f / sample_rate;
    st->dt_fdtd = (float)amr_get_dt(st->amr);
    if (st->dt_fdtd <= 0.0f) return fail("dt_fdtd_invalid");
    st->modal_stride = modal_stride;
    st->K = (int)ceilf(st->dt_audio / st->dt_fdtd);
    if (st->K < 1) st->K = 1;
    if (st->K > kMaxSub) st->K = kMaxSub;
    st->nb = d->n_bridge_plate;                                          // :1459-1477
    st->bidx.assign(d->bridge_plate_idx, d->bridge_plate_idx + st->nb);
    st->bwgt.assign(d->bridge_plate_wgt, d->bridge_plate_wgt + st->nb);
    if (d->n_neck_plate > 0 && d->neck_plate_idx && d->neck_plate_wgt) {
        st->nn = d->n_neck_plate;
        st->nidx.assign(d->neck_plate_idx, d->neck_plate_idx + st->nn);
        st->nwgt.assign(d->neck_plate_wgt, d->neck_plate_wgt + st->nn);
    }
    const float sigma = (d->plate_dx > 0.0f) ? 2.0f * d->plate_dx : 0.005f;   // :1711-1712
    if (const char* stage = co_setup(st, n_strings, sd, n_pickups, pd, n_mics, md, nullptr, sigma))
        return fail(stage);
    return st;
}

SK_API const char* coevolver_create_amr_last_error(void) { return g_amr_err.c_str(); }

// Extension: the pressure backend's error code name and message of the last
// failed coevolver_create_amr on this thread (the reference's pybind layer
// reads them through C++ calls, pybind_kernels.cpp:6060-6080).
SK_API const char* coevolver_amr_backend_error_code(void) {
    return amr_pressure_backend_last_error_code();
}
SK_API const char* coevolver_amr_backend_error_message(void) {
    return amr_pressure_backend_last_error_message();
}

SK_API void coevolver_destroy(AcousticCoEvolverState* st) {
    if (!st) return;
    if (st->worker.joinable()) {                                         // :1800-1807
        st->cancel.store(1);
        st->worker.join();
    }
    co_free(st);
    delete st;
}

SK_API int coevolver_pluck_string(AcousticCoEvolverState* st, int string_idx,
                                  float position_norm, float amplitude, int n_duration_samples) {
    if (!st) return CE_ERR_NULL;
    if (busy(st)) return CE_ERR_BUSY;
    if (string_idx < 0 || string_idx >= st->ns) return CE_ERR_STRING_IDX;
    return pluck_host(st, string_idx, position_norm, amplitude, n_duration_samples);
}

SK_API int coevolver_inject_string_force(AcousticCoEvolverState* st, int string_idx,
                                         int n_samples, const float* force_buf,
                                         float force_pos_norm) {             // :2472-2509
    if (!st || !force_buf) return CE_ERR_NULL;
    if (busy(st)) return CE_ERR_BUSY;
    if (string_idx < 0 || string_idx >= st->ns) return CE_ERR_STRING_IDX;
    if (n_samples <= 0) return CE_OK;
    HString& s = st->s[string_idx];
    const int N = s.n_nodes;
    const float pos = fmaxf(0.0f, fminf(1.0f, force_pos_norm));
    int i0 = (int)roundf(pos * (float)(N - 1));
    if (i0 <= 0) i0 = 1;
    if (i0 >= N - 1) i0 = N - 2;
    const float sc = s.dt_s * s.dt_s / (s.mu * s.ds);
    s.fbuf.resize(n_samples);
    for (int q = 0; q < n_samples; ++q) s.fbuf[q] = force_buf[q] * sc;
    s.fnode = i0;
    s.fonset = st->written;
    return CE_OK;
}

SK_API int coevolver_schedule_pluck(AcousticCoEvolverState* st, int onset_sample, int string_idx,
                                    float position_norm, float amplitude,
                                    int n_duration_samples) {                // :1756-1786
    if (!st) return CE_ERR_NULL;
    if (busy(st)) return CE_ERR_BUSY;
    if (string_idx < 0 || string_idx >= st->ns) return CE_ERR_PARAM;
    if (onset_sample < 0) onset_sample = 0;
    HPluck e{onset_sample, string_idx, position_norm, amplitude, n_duration_samples};
    size_t q = st->ev.size();
    st->ev.push_back(e);
    while (q > 0 && st->ev[q - 1].onset > onset_sample) {
        st->ev[q] = st->ev[q - 1];
        --q;
    }
    st->ev[q] = e;
    return CE_OK;
}

SK_API void coevolver_clear_pluck_schedule(AcousticCoEvolverState* st) {
    if (!st || busy(st)) return;
    st->ev.clear();
    st->cursor = 0;
}

SK_API int coevolver_step(AcousticCoEvolverState* st, int n_samples) {
    if (!st) return CE_ERR_NULL;
    if (busy(st)) return CE_ERR_BUSY;
    if (n_samples <= 0) return CE_OK;
    return step_samples(st, n_samples, kStep, nullptr, 0.0f, nullptr, 0, nullptr, nullptr);
}

SK_API int coevolver_step_timed(AcousticCoEvolverState* st, int n_samples, float* ms_total) {
    if (!st || !ms_total) return CE_ERR_NULL;
    if (busy(st)) return CE_ERR_BUSY;
    *ms_total = 0.0f;
    if (n_samples <= 0) return CE_OK;
    return step_samples(st, n_samples, kStep, nullptr, 0.0f, nullptr, 0, ms_total, nullptr);
}

SK_API int coevolver_step_block_with_drive(AcousticCoEvolverState* st,
                                           const float* const* string_drive_blocks,
                                           int n_strings, int n_samples, float force_pos_norm,
                                           float* mic_out_buf) {       // :2253-2321
    if (!st || !mic_out_buf) return CE_ERR_NULL;
    if (busy(st)) return CE_ERR_BUSY;
    if (n_strings != st->ns) return CE_ERR_PARAM;
    if (n_samples < 0) return CE_ERR_DIM;
    if (n_samples == 0) return CE_OK;
    if (st->nmic < 1) return CE_ERR_DIM;
    return step_samples(st, n_samples, kDrive, string_drive_blocks, force_pos_norm, mic_out_buf,
                        0, nullptr, nullptr);
}

SK_API int coevolver_step_async(AcousticCoEvolverState* st, int n_samples) {   // :2355-2374
    if (!st) return CE_ERR_NULL;
    if (n_samples <= 0) return CE_OK;
    if (st->status.load() == CE_STATUS_RUNNING) return CE_ERR_PARAM;
    if (st->worker.joinable()) st->worker.join();
    st->cancel.store(0);
    st->progress.store(0);
    st->total.store(n_samples);
    st->last_error.store(CE_OK);
    st->status.store(CE_STATUS_RUNNING, std::memory_order_release);
    st->worker = std::thread([st, n_samples]() {
        // progress per chunk of 64 samples; cancellation between chunks
        int off = 0;
        while (off < n_samples) {
            if (st->cancel.load(std::memory_order_acquire)) {
                st->status.store(CE_STATUS_CANCELLED, std::memory_order_release);
                return;
            }
            const int m = std::min(64, n_samples - off);
            int done = 0;
            const int rc = step_samples(st, m, kWorker, nullptr, 0.0f, nullptr, off, nullptr, &done);
            if (rc != CE_OK) {
                st->progress.fetch_add(done + 1);
                st->last_error.store(rc);
                st->status.store(CE_STATUS_ERROR, std::memory_order_release);
                return;
            }
            st->progress.fetch_add(m);
            off += m;
        }
        st->status.store(CE_STATUS_DONE, std::memory_order_release);
    });
    return CE_OK;
}

SK_API void coevolver_get_progress(const AcousticCoEvolverState* st, CoEvolverProgress* out) {
    if (!st || !out) return;
    out->samples_done = st->progress.load();
    out->samples_total = st->total.load();
    out->status = st->status.load(std::memory_order_acquire);
    out->error_code = st->last_error.load();
}

SK_API int coevolver_wait(AcousticCoEvolverState* st) {
    if (!st) return CE_ERR_NULL;
    if (st->worker.joinable()) st->worker.join();
    return st->last_error.load();
}

SK_API int coevolver_cancel(AcousticCoEvolverState* st) {
    if (!st) return CE_ERR_NULL;
    st->cancel.store(1, std::memory_order_release);
    if (st->worker.joinable()) st->worker.join();
    return CE_OK;
}

SK_API int coevolver_is_running(const AcousticCoEvolverState* st) {
    if (!st) return 0;
    return busy(st) ? 1 : 0;
}

SK_API int coevolver_get_pickup_output(const AcousticCoEvolverState* cst, int pickup_idx,
                                       float* out, int n_samples) {
    if (!cst || !out) return CE_ERR_NULL;
    if (pickup_idx < 0 || pickup_idx >= cst->npk) return CE_ERR_DIM;
    AcousticCoEvolverState* st = const_cast<AcousticCoEvolverState*>(cst);
    return copy_ring(st, st->d_pk_out + (size_t)pickup_idx * kRing, out, n_samples);
}

SK_API int coevolver_get_mic_output(const AcousticCoEvolverState* cst, int mic_idx, float* out,
                                    int n_samples) {
    if (!cst || !out) return CE_ERR_NULL;
    if (mic_idx < 0 || mic_idx >= cst->nmic) return CE_ERR_DIM;
    AcousticCoEvolverState* st = const_cast<AcousticCoEvolverState*>(cst);
    return copy_ring(st, st->d_mic_out + (size_t)mic_idx * kRing, out, n_samples);
}

SK_API int coevolver_get_pressure_field(const AcousticCoEvolverState* st, float* out,
                                        int out_len) {
    if (!st) return CE_ERR_NULL;
    if (st->amr) return amr_get_pressure(st->amr, out, out_len);        // flat cell array
    return fdtd_get_pressure_field(st->fdtd, out, out_len);
}

SK_API int coevolver_get_pressure_field_uniform(const AcousticCoEvolverState* st, int Nx,
                                                int Ny, int Nz, const double bmin[3],
                                                const double bmax[3], float* out,
                                                int out_len) {
    if (!st || !out || !bmin || !bmax) return CE_ERR_NULL;
    if (out_len != Nx * Ny * Nz) return CE_ERR_DIM;
    if (st->amr) return amr_scatter_pressure_uniform(st->amr, Nx, Ny, Nz, bmin, bmax, out);
    return fdtd_get_pressure_field(st->fdtd, out, out_len);
}

SK_API int coevolver_get_amr_cell_centers(const AcousticCoEvolverState* st, float* out_xyz,
                                          int n_cells_3) {
    if (!st || !out_xyz) return CE_ERR_NULL;
    if (!st->amr) return CE_ERR_PARAM;                                   // uniform body
    return amr_get_cell_centers(st->amr, out_xyz, n_cells_3);
}

SK_API int coevolver_get_plate_displacement(const AcousticCoEvolverState* st, float* out,
                                            int out_len) {
    if (!st) return CE_ERR_NULL;
    if (st->amr) return amr_get_plate_displacement(st->amr, out, out_len);
    return fdtd_get_plate_displacement(st->fdtd, out, out_len);
}

SK_API int coevolver_get_plate_dims(const AcousticCoEvolverState* st, int* out_Nx, int* out_Ny,
                                    int* out_count) {
    if (!st || !out_Nx || !out_Ny || !out_count) return CE_ERR_NULL;
    *out_Nx = st->Nx;
    *out_Ny = st->Ny;
    *out_count = st->Nx * st->Ny;
    return CE_OK;
}

SK_API int coevolver_get_modal_amplitudes(const AcousticCoEvolverState* st, int string_idx,
                                          float* out_re, float* out_im, int n_modes) {
    if (!st || !out_re || !out_im) return CE_ERR_NULL;
    if (string_idx < 0 || string_idx >= st->ns) return CE_ERR_STRING_IDX;
    const int m = n_modes < kModes ? n_modes : kModes;
    float h[2 * kModes];
    if (hipMemcpy(h, st->d_modal + (size_t)string_idx * 2 * kModes, sizeof(h),
                  hipMemcpyDeviceToHost) != hipSuccess)
        return CE_ERR_NULL;
    for (int k = 0; k < m; ++k) {
        out_re[k] = h[k];
        out_im[k] = h[kModes + k];
    }
    return CE_OK;
}

SK_API int coevolver_sample_pressure(const AcousticCoEvolverState* st, int n_rec,
                                     const float* rec_xyz, float* out_p) {   // :2608-2625
    if (!st || !rec_xyz || !out_p) return CE_ERR_NULL;
    if (st->amr) return CE_ERR_PARAM;                                    // :2613-2616
    std::vector<float> g((size_t)3 * std::max(n_rec, 0));
    for (int r = 0; r < n_rec; ++r)
        for (int q = 0; q < 3; ++q) g[3 * r + q] = (rec_xyz[3 * r + q] - st->origin[q]) / st->dx;
    const int rc = fdtd_sample_pressure(st->fdtd, n_rec, g.data(), out_p);
    return rc == FDTD_OK ? CE_OK : CE_ERR_NULL;
}

SK_API int coevolver_get_string_velocity(const AcousticCoEvolverState* st, int string_idx,
                                         float* out_v, int out_len) {        // :2627-2646
    if (!st || !out_v) return CE_ERR_NULL;
    if (string_idx < 0 || string_idx >= st->ns) return CE_ERR_STRING_IDX;
    const HString& s = st->s[string_idx];
    if (out_len != s.n_segs * 3) return CE_ERR_DIM;
    std::vector<float> u;
    if (read_strings(st, u) != CE_OK) return CE_ERR_NULL;
    const float inv = 1.0f / s.dt_s;
    const size_t T = st->tot_nodes;
    for (int q = 0; q < s.n_segs; ++q) {
        const size_t n = s.node_off + q;
        const float v0 = (u[n] - u[2 * T + n]) * inv;
        const float v1 = (u[T + n] - u[3 * T + n]) * inv;
        const float* nn = &s.nrm[q * 3];
        const float* b = &s.bin[q * 3];
        for (int k = 0; k < 3; ++k) out_v[q * 3 + k] = v0 * nn[k] + v1 * b[k];
    }
    return CE_OK;
}

SK_API int coevolver_get_string_displacement(const AcousticCoEvolverState* st, int string_idx,
                                             float* out_d, int out_len) {    // :2648-2666
    if (!st || !out_d) return CE_ERR_NULL;
    if (string_idx < 0 || string_idx >= st->ns) return CE_ERR_STRING_IDX;
    const HString& s = st->s[string_idx];
    if (out_len != s.n_segs * 3) return CE_ERR_DIM;
    std::vector<float> u;
    if (read_strings(st, u) != CE_OK) return CE_ERR_NULL;
    const size_t T = st->tot_nodes;
    for (int q = 0; q < s.n_segs; ++q) {
        const size_t n = s.node_off + q;
        const float d0 = u[n], d1 = u[T + n];
        const float* nn = &s.nrm[q * 3];
        const float* b = &s.bin[q * 3];
        for (int k = 0; k < 3; ++k) out_d[q * 3 + k] = d0 * nn[k] + d1 * b[k];
    }
    return CE_OK;
}

SK_API int coevolver_get_string_position(const AcousticCoEvolverState* st, int string_idx,
                                         float* out_xyz, int out_len) {      // :2668-2690
    if (!st || !out_xyz) return CE_ERR_NULL;
    if (string_idx < 0 || string_idx >= st->ns) return CE_ERR_STRING_IDX;
    const HString& s = st->s[string_idx];
    if (out_len != s.n_nodes * 3) return CE_ERR_DIM;
    std::vector<float> u;
    if (read_strings(st, u) != CE_OK) return CE_ERR_NULL;
    const size_t T = st->tot_nodes;
    for (int nd = 0; nd < s.n_nodes; ++nd) {
        const int q = (nd < s.n_segs) ? nd : s.n_segs - 1;
        const size_t n = s.node_off + nd;
        const float d0 = u[n], d1 = u[T + n];
        const float* base = &s.xyz[nd * 3];
        const float* nn = &s.nrm[q * 3];
        const float* b = &s.bin[q * 3];
        for (int k = 0; k < 3; ++k) out_xyz[nd * 3 + k] = base[k] + d0 * nn[k] + d1 * b[k];
    }
    return CE_OK;
}

SK_API int coevolver_get_n_strings(const AcousticCoEvolverState* st) { return st ? st->ns : 0; }
SK_API int coevolver_get_n_pickups(const AcousticCoEvolverState* st) { return st ? st->npk : 0; }
SK_API int coevolver_get_n_mics(const AcousticCoEvolverState* st) { return st ? st->nmic : 0; }
SK_API float coevolver_get_dt_audio(const AcousticCoEvolverState* st) {
    return st ? st->dt_audio : 0.0f;
}
SK_API float coevolver_get_dt_fdtd(const AcousticCoEvolverState* st) {
    return st ? st->dt_fdtd : 0.0f;
}

SK_API int coevolver_reset(AcousticCoEvolverState* st) {                  // :2698-2738
    if (!st) return CE_ERR_NULL;
    if (busy(st)) return CE_ERR_BUSY;
    // the body: a uniform FDTD is zeroed; an AMR body is left as it is (the
    // reference calls fdtd_reset on its null uniform handle, :2702)
    if (st->fdtd && fdtd_reset(st->fdtd) != FDTD_OK) return CE_ERR_NULL;
    hipStream_t sm = co_stream(st);
    std::vector<float> sw((size_t)kSw * std::max(st->ns, 1));
    if (st->ns && hipMemcpy(sw.data(), st->d_sw, sw.size() * sizeof(float),
                            hipMemcpyDeviceToHost) != hipSuccess)
        return CE_ERR_NULL;
    for (int i = 0; i < st->ns; ++i) {
        float* w = &sw[(size_t)i * kSw];
        const float teff = w[kTeff];                 // the effective tension is kept
        for (int q = 0; q < kSw; ++q) w[q] = 0.0f;
        w[kTeff] = teff;
        st->s[i].fbuf.clear();
    }
    if ((st->ns && hipMemcpy(st->d_sw, sw.data(), sw.size() * sizeof(float),
                             hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemsetAsync(st->d_u, 0, sizeof(float) * 4 * std::max(st->tot_nodes, 1), sm) != hipSuccess ||
        hipMemsetAsync(st->d_modal, 0, sizeof(float) * 2 * kModes * std::max(st->ns, 1), sm) != hipSuccess ||
        hipMemsetAsync(st->d_neck, 0, sizeof(float) * 4, sm) != hipSuccess ||
        hipMemsetAsync(st->d_pk_out, 0, sizeof(float) * kRing * std::max(st->npk, 1), sm) != hipSuccess ||
        hipMemsetAsync(st->d_mic_out, 0, sizeof(float) * kRing * std::max(st->nmic, 1), sm) != hipSuccess ||
        hipStreamSynchronize(sm) != hipSuccess)
        return CE_ERR_NULL;
    st->gamma_scale = 1.0f;
    st->out_pos = 0;
    st->written = 0;
    st->modal_counter = 0;
    st->cursor = 0;
    return CE_OK;
}

SK_API int coevolver_set_face_fractions(AcousticCoEvolverState* st, const float* vx_frac,
                                        const float* vy_frac, const float* vz_frac) {
    if (!st || !st->fdtd) return CE_ERR_NULL;
    return fdtd_set_face_fractions(st->fdtd, vx_frac, vy_frac, vz_frac);
}

SK_API void coevolver_set_damping_scale(AcousticCoEvolverState* st, float scale) {
    if (!st || busy(st)) return;
    st->gamma_scale = (scale > 0.0f) ? scale : 1.0f;
}

SK_API int coevolver_get_surface_emission(const AcousticCoEvolverState* st, int n_surf,
                                          const float* surf_xyz, const float* surf_normals,
                                          float* out_P, float* out_vn) {    // :2760-2779
    if (!st || !surf_xyz || !surf_normals || !out_P || !out_vn) return CE_ERR_NULL;
    if (!st->fdtd) return CE_ERR_NULL;                                   // no uniform grid
    std::vector<float> g((size_t)3 * std::max(n_surf, 0));
    for (int r = 0; r < n_surf; ++r)
        for (int q = 0; q < 3; ++q) g[3 * r + q] = (surf_xyz[3 * r + q] - st->origin[q]) / st->dx;
    const int rc = fdtd_get_surface_emission(st->fdtd, n_surf, g.data(), surf_normals, out_P,
                                             out_vn);
    return rc == FDTD_OK ? CE_OK : CE_ERR_NULL;
}

SK_API int coevolver_sample_velocity(const AcousticCoEvolverState* st, int n_rec,
                                     const float* rec_xyz, float* out_vx, float* out_vy,
                                     float* out_vz) {                        // :2781-2797
    if (!st || !rec_xyz || !out_vx || !out_vy || !out_vz) return CE_ERR_NULL;
    if (!st->fdtd) return CE_ERR_NULL;                                   // no uniform grid
    std::vector<float> g((size_t)3 * std::max(n_rec, 0));
    for (int r = 0; r < n_rec; ++r)
        for (int q = 0; q < 3; ++q) g[3 * r + q] = (rec_xyz[3 * r + q] - st->origin[q]) / st->dx;
    const int rc = fdtd_sample_velocity(st->fdtd, n_rec, g.data(), out_vx, out_vy, out_vz);
    return rc == FDTD_OK ? CE_OK : CE_ERR_NULL;
}

SK_API AcousticFDTDState* coevolver_get_fdtd(AcousticCoEvolverState* st) {
    return st ? st->fdtd : nullptr;
}

SK_API int coevolver_get_substeps(const AcousticCoEvolverState* st) { return st ? st->K : 0; }

SK_API int coevolver_get_string_substeps(const AcousticCoEvolverState* st, int string_idx) {
    if (!st || string_idx < 0 || string_idx >= st->ns) return -1;
    return st->s[string_idx].n_sub;
}

SK_API int coevolver_get_samples_written(const AcousticCoEvolverState* st) {
    return st ? st->written : 0;
}

SK_API int coevolver_get_string_state(const AcousticCoEvolverState* st, int string_idx,
                                      float* out, int out_len) {
    if (!st || !out) return CE_ERR_NULL;
    if (string_idx < 0 || string_idx >= st->ns) return CE_ERR_STRING_IDX;
    const HString& s = st->s[string_idx];
    const int N = s.n_nodes;
    if (out_len != 4 * N + 6) return CE_ERR_DIM;
    std::vector<float> u;
    if (read_strings(st, u) != CE_OK) return CE_ERR_NULL;
    float w[kSw];
    if (hipMemcpy(w, st->d_sw + (size_t)string_idx * kSw, sizeof(w), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return CE_ERR_NULL;
    const size_t T = st->tot_nodes;
    for (int b = 0; b < 4; ++b) {
        // u0, u1, up0, up1 from the device order u0, u1, up0, up1
        const size_t src = (size_t)b * T + s.node_off;
        for (int i = 0; i < N; ++i) out[b * N + i] = u[src + i];
    }
    out[4 * N + 0] = w[kTeff];
    out[4 * N + 1] = w[kSad];
    out[4 * N + 2] = w[kSadP];
    out[4 * N + 3] = w[kBr];
    out[4 * N + 4] = w[kBrP];
    out[4 * N + 5] = (float)s.n_sub;
    return CE_OK;
}

SK_API int coevolver_get_neck_state(const AcousticCoEvolverState* st, float* out3) {
    if (!st || !out3) return CE_ERR_NULL;
    if (hipMemcpy(out3, st->d_neck, 3 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return CE_ERR_NULL;
    return CE_OK;
}

}  // extern "C"
