// afdtd_api.hip — host side of libafdtd.so: the reference's 26 fdtd_*
// functions (csrc/include/acoustic_fdtd.h) and the extensions of afdtd.h.
//
// Set-up constants (dt, damping multipliers, plate coefficients) are computed
// here on the host with the reference's float expressions, so the device
// never re-derives them.  All device work of a handle runs on its own HIP
// stream; the getters are synchronous on return.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "afdtd_internal.h"

static thread_local std::string g_last_error;
void afdtd_set_error(const std::string& msg) { g_last_error = msg; }
extern "C" SK_API const char* fdtd_last_error(void) { return g_last_error.c_str(); }

// ── Scratch of the allocation-free sampler ───────────────────────────────
// fdtd_sample_pressure_precomputed_noalloc stages the 8-tap stencils and the
// outputs of up to FDTD_NOALLOC_MAX_RECEIVERS receivers here instead of in
// buffers the handle allocates: 8 int32 indices, 8 float weights and one
// float output per receiver (4.25 MiB).  Every GPU holds its own copy for the
// life of the process.  In host code `scratch` names a host-side stand-in,
// not a device address: host code uses only the per-device address that
// hipGetSymbolAddress returns (resolve_scratch), never the array itself.
namespace {
constexpr size_t kScrRec = FDTD_NOALLOC_MAX_RECEIVERS;
constexpr size_t kScrIdx = 0;                                  // int32[8 * kScrRec]
constexpr size_t kScrWgt = kScrIdx + 8 * sizeof(int) * kScrRec;  // float[8 * kScrRec]
constexpr size_t kScrOut = kScrWgt + 8 * sizeof(float) * kScrRec;  // float[kScrRec]
constexpr size_t kScrBytes = kScrOut + sizeof(float) * kScrRec;
constexpr int kScrMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kScrMaxDevices] = {};
}  // namespace
static __device__ __align__(256) unsigned char scratch[kScrBytes];

// Address of the current device's copy of `scratch`, looked up once per
// device (with that device current) and kept for the process.
static int resolve_scratch(unsigned char** out) {
    *out = nullptr;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScrMaxDevices) {
        afdtd_set_error("allocation-free scratch: no current HIP device");
        return FDTD_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            afdtd_set_error(std::string("allocation-free scratch: hipGetSymbolAddress: ") +
                            hipGetErrorString(e));
            return FDTD_ERR_DEVICE;
        }
        g_scratch_dev[dev] = (unsigned char*)ptr;
    }
    *out = g_scratch_dev[dev];
    return FDTD_OK;
}

namespace {

inline int is_solid_h(uint8_t c) { return c == FDTD_WALL || c == FDTD_PLATE; }

// PML profile (acoustic_fdtd.cpp:216-222), float, same expression order.
float pml_sigma(int d, int n_pml, float dt) {
    if (n_pml <= 0 || d <= 0) return 0.0f;
    float frac = (float)d / (float)n_pml;
    float sigma_max = 3.45f / (dt * (float)n_pml);
    return sigma_max * frac * frac * frac;
}

int pml_depth_h(int i, int n, int np) {
    return (i < np) ? (np - i) : ((i >= n - np) ? (i - (n - np) + 1) : 0);
}

template <typename T>
hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

int storage_bytes(int storage) { return storage == FDTD_STORE_F32 ? 4 : 2; }
// bytes of one field buffer: N elements, plus (at the next 4-byte boundary)
// one int32 exponent per 64-cell z-row segment for block floating point
size_t field_bytes(int storage, int Nx, int Ny, int Nz) {
    const size_t N = (size_t)Nx * Ny * Nz;
    size_t b = N * storage_bytes(storage);
    if (storage == FDTD_STORE_BFP16 || storage == FDTD_STORE_F16S)
        b = ((N * 2 + 3) & ~(size_t)3) + sizeof(int) * (size_t)Nx * Ny * ((Nz + 63) / 64);
    return b;
}

// Copy a packed staggered array (reference layout) into the padded layout.
// Packed rows of `w` elements at pitch `w`, padded rows at pitch `pitch`.
hipError_t upload_rows(float* dst, const float* src, size_t w, size_t pitch, size_t rows,
                       hipStream_t s) {
    if (w == 0 || rows == 0) return hipSuccess;
    return hipMemcpy2DAsync(dst, pitch * sizeof(float), src, w * sizeof(float),
                            w * sizeof(float), rows, hipMemcpyHostToDevice, s);
}

int sync(AcousticFDTDState* st) {
    AFDTD_CHECK(hipStreamSynchronize(st->stream));
    return FDTD_OK;
}

// Current pressure / a velocity component as FP32 device memory (converting
// 16-bit storage into the staging buffer).
const float* field_f32(const AcousticFDTDState* st, int which /*0=P,1..3=V*/) {
    const void* src = which == 0 ? st->P[st->cur] : st->V[which - 1];
    if (st->storage == FDTD_STORE_F32) return (const float*)src;
    float scale = which == 0 ? 1.0f : st->sp.inv_sV;
    (void)launch_to_f32(src, st->storage, scale, st->stage, (size_t)st->N, st->stream, st->Nz);
    return st->stage;
}

int ensure_gather(AcousticFDTDState* st, int n) {
    if (n <= st->gcap) return FDTD_OK;
    dfree(st->d_gidx); dfree(st->d_gwgt); dfree(st->d_gout);
    st->d_gidx = nullptr; st->d_gwgt = nullptr; st->d_gout = nullptr; st->gcap = 0;
    AFDTD_CHECK(dalloc(&st->d_gidx, (size_t)n * 8));
    AFDTD_CHECK(dalloc(&st->d_gwgt, (size_t)n * 8));
    AFDTD_CHECK(dalloc(&st->d_gout, (size_t)n));
    st->gcap = n;
    return FDTD_OK;
}

// Gather n outputs of `taps` weighted samples of a field (device), in either
// summation order, into host out.
// Every stencil index must address the gathered field (device memory is not
// bounds-checked): [0, N) for the pressure and the padded velocity layouts,
// [0, Nplate) for the plate.
int check_indices(const AcousticFDTDState* st, int which, const int* idx, size_t count) {
    const int lim = which == 4 ? st->Nplate : st->N;
    for (size_t q = 0; q < count; ++q)
        if (idx[q] < 0 || idx[q] >= lim) {
            afdtd_set_error("sampler index " + std::to_string(idx[q]) + " outside [0, " +
                            std::to_string(lim) + ")");
            return FDTD_ERR_DIM;
        }
    return FDTD_OK;
}

int gather_host(AcousticFDTDState* st, int which, const int* idx, const float* wgt, int n,
                int taps, int zero_start, float* out) {
    if (n <= 0) return FDTD_OK;
    int rc = check_indices(st, which, idx, (size_t)n * taps);
    if (rc) return rc;
    rc = ensure_gather(st, n);       // room for n outputs of up to 8 taps
    if (rc) return rc;
    AFDTD_CHECK(hipMemcpyAsync(st->d_gidx, idx, sizeof(int) * n * taps, hipMemcpyHostToDevice,
                               st->stream));
    AFDTD_CHECK(hipMemcpyAsync(st->d_gwgt, wgt, sizeof(float) * n * taps, hipMemcpyHostToDevice,
                               st->stream));
    const void* F = which == 0 ? st->P[st->cur] : st->V[which - 1];
    float scale = which == 0 ? 1.0f : st->sp.inv_sV;
    if (which == 4) { F = st->d_w[0]; scale = 1.0f; }
    int storage = which == 4 ? FDTD_STORE_F32 : st->storage;
    AFDTD_CHECK(launch_gather(F, storage, scale, st->d_gidx, st->d_gwgt, n, taps, zero_start,
                              st->d_gout, st->stream, st->N, st->Nz));
    AFDTD_CHECK(hipMemcpyAsync(out, st->d_gout, sizeof(float) * n, hipMemcpyDeviceToHost,
                               st->stream));
    return sync(st);
}

// Trilinear 8-tap pressure stencil (acoustic_fdtd.cpp:1575-1594).
void p_sampler(int Nx, int Ny, int Nz, const float* xyz, int* idx, float* wgt) {
    float gx = xyz[0], gy = xyz[1], gz = xyz[2];
    if (gx < 0.0f) gx = 0.0f;
    if (gx > (float)(Nx - 1)) gx = (float)(Nx - 1);
    if (gy < 0.0f) gy = 0.0f;
    if (gy > (float)(Ny - 1)) gy = (float)(Ny - 1);
    if (gz < 0.0f) gz = 0.0f;
    if (gz > (float)(Nz - 1)) gz = (float)(Nz - 1);
    int i0 = (int)gx, i1 = (i0 < Nx - 1) ? i0 + 1 : i0;
    int j0 = (int)gy, j1 = (j0 < Ny - 1) ? j0 + 1 : j0;
    int k0 = (int)gz, k1 = (k0 < Nz - 1) ? k0 + 1 : k0;
    float fx = gx - (float)i0, fy = gy - (float)j0, fz = gz - (float)k0;
    auto I = [&](int i, int j, int k) { return k + Nz * (j + Ny * i); };
    idx[0] = I(i0, j0, k0); wgt[0] = (1 - fx) * (1 - fy) * (1 - fz);
    idx[1] = I(i1, j0, k0); wgt[1] = (fx) * (1 - fy) * (1 - fz);
    idx[2] = I(i0, j1, k0); wgt[2] = (1 - fx) * (fy) * (1 - fz);
    idx[3] = I(i1, j1, k0); wgt[3] = (fx) * (fy) * (1 - fz);
    idx[4] = I(i0, j0, k1); wgt[4] = (1 - fx) * (1 - fy) * (fz);
    idx[5] = I(i1, j0, k1); wgt[5] = (fx) * (1 - fy) * (fz);
    idx[6] = I(i0, j1, k1); wgt[6] = (1 - fx) * (fy) * (fz);
    idx[7] = I(i1, j1, k1); wgt[7] = (fx) * (fy) * (fz);
}

// Staggered 8-tap stencil for velocity component `ax` with the half-cell
// shift (acoustic_fdtd.cpp:1601-1645); indices in the PACKED layout.
void v_sampler(int Nx, int Ny, int Nz, int ax, const float* xyz, int* idx, float* wgt) {
    int Nx1 = ax == 0 ? Nx - 1 : Nx, Ny1 = ax == 1 ? Ny - 1 : Ny, Nz1 = ax == 2 ? Nz - 1 : Nz;
    float gx = xyz[0] - (ax == 0 ? 0.5f : 0.0f);
    float gy = xyz[1] - (ax == 1 ? 0.5f : 0.0f);
    float gz = xyz[2] - (ax == 2 ? 0.5f : 0.0f);
    if (gx < 0.0f) gx = 0.0f;
    if (gx > (float)(Nx1 - 1)) gx = (float)(Nx1 - 1);
    if (gy < 0.0f) gy = 0.0f;
    if (gy > (float)(Ny1 - 1)) gy = (float)(Ny1 - 1);
    if (gz < 0.0f) gz = 0.0f;
    if (gz > (float)(Nz1 - 1)) gz = (float)(Nz1 - 1);
    int i0 = (int)gx, i1 = (i0 < Nx1 - 1) ? i0 + 1 : i0;
    int j0 = (int)gy, j1 = (j0 < Ny1 - 1) ? j0 + 1 : j0;
    int k0 = (int)gz, k1 = (k0 < Nz1 - 1) ? k0 + 1 : k0;
    float fx = gx - (float)i0, fy = gy - (float)j0, fz = gz - (float)k0;
    auto I = [&](int a, int b, int c) {
        return ax == 2 ? (c + Nz1 * (b + Ny * a))
                       : (ax == 1 ? (c + Nz * (b + Ny1 * a)) : (c + Nz * (b + Ny * a)));
    };
    idx[0] = I(i0, j0, k0); wgt[0] = (1 - fx) * (1 - fy) * (1 - fz);
    idx[1] = I(i1, j0, k0); wgt[1] = (fx) * (1 - fy) * (1 - fz);
    idx[2] = I(i0, j1, k0); wgt[2] = (1 - fx) * (fy) * (1 - fz);
    idx[3] = I(i1, j1, k0); wgt[3] = (fx) * (fy) * (1 - fz);
    idx[4] = I(i0, j0, k1); wgt[4] = (1 - fx) * (1 - fy) * (fz);
    idx[5] = I(i1, j0, k1); wgt[5] = (fx) * (1 - fy) * (fz);
    idx[6] = I(i0, j1, k1); wgt[6] = (1 - fx) * (fy) * (fz);
    idx[7] = I(i1, j1, k1); wgt[7] = (fx) * (fy) * (fz);
}

// Packed staggered index -> padded device index.
inline int packed_to_padded(int ax, int q, int Nx, int Ny, int Nz) {
    (void)Nx;
    if (ax == 0) return q;                         // Vx packed == padded prefix
    if (ax == 1) {                                 // k + Nz*(j + (Ny-1)*i)
        int k = q % Nz, t = q / Nz, j = t % (Ny - 1), i = t / (Ny - 1);
        return k + Nz * (j + Ny * i);
    }
    int k = q % (Nz - 1), t = q / (Nz - 1);        // k + (Nz-1)*(j + Ny*i)
    return k + Nz * t;
}

// Entries of a force injection grouped by target plate node p = idx/div
// (entries out of range or on inactive nodes dropped, as the reference's
// loops skip them): start[ng+1] | node[ng] | entry[...], each node's entries
// in call order.  Returns the group count.
int force_groups(const int* idx, int n, int div, const uint8_t* mask, int nmax,
                 std::vector<int>& grp) {
    std::vector<std::pair<int, int>> e;
    e.reserve(n);
    for (int i = 0; i < n; ++i) {
        const int p = idx[i] / div;
        if (p < 0 || p >= nmax) continue;
        if (mask && !mask[p]) continue;
        e.emplace_back(p, i);
    }
    std::stable_sort(e.begin(), e.end(),
                     [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                         return a.first < b.first;
                     });
    std::vector<int> start, node;
    for (size_t q = 0; q < e.size(); ++q)
        if (q == 0 || e[q].first != e[q - 1].first) {
            start.push_back((int)q);
            node.push_back(e[q].first);
        }
    start.push_back((int)e.size());
    const int ng = (int)node.size();
    grp.clear();
    grp.insert(grp.end(), start.begin(), start.end());
    grp.insert(grp.end(), node.begin(), node.end());
    for (const auto& x : e) grp.push_back(x.second);
    return ng;
}

int upload_fracs(AcousticFDTDState* st, const float* fx, const float* fy, const float* fz) {
    int Nx = st->Nx, Ny = st->Ny, Nz = st->Nz;
    if (fx && st->Nvx > 0)
        AFDTD_CHECK(hipMemcpyAsync(st->d_fx, fx, sizeof(float) * st->Nvx, hipMemcpyHostToDevice,
                                   st->stream));
    if (fy && st->Nvy > 0)
        AFDTD_CHECK(upload_rows(st->d_fy, fy, (size_t)(Ny - 1) * Nz, (size_t)Ny * Nz, Nx,
                                st->stream));
    if (fz && st->Nvz > 0)
        AFDTD_CHECK(upload_rows(st->d_fz, fz, (size_t)(Nz - 1), (size_t)Nz, (size_t)Nx * Ny,
                                st->stream));
    return sync(st);
}

// Default fractions (acoustic_fdtd.cpp:497-534) in packed layouts.
void default_fracs(const AcousticFDTDState* st, std::vector<float>& fx, std::vector<float>& fy,
                   std::vector<float>& fz) {
    int Nx = st->Nx, Ny = st->Ny, Nz = st->Nz, piz = st->plate_iz;
    const uint8_t* ct = st->h_ct.data();
    auto I = [&](int i, int j, int k) { return k + Nz * (j + Ny * i); };
    fx.assign(std::max(st->Nvx, 0), 1.0f);
    fy.assign(std::max(st->Nvy, 0), 1.0f);
    fz.assign(std::max(st->Nvz, 0), 1.0f);
    for (int i = 0; i < Nx - 1; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k)
                fx[k + Nz * (j + Ny * i)] =
                    (is_solid_h(ct[I(i, j, k)]) || is_solid_h(ct[I(i + 1, j, k)])) ? 0.0f : 1.0f;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny - 1; ++j)
            for (int k = 0; k < Nz; ++k)
                fy[k + Nz * (j + (Ny - 1) * i)] =
                    (is_solid_h(ct[I(i, j, k)]) || is_solid_h(ct[I(i, j + 1, k)])) ? 0.0f : 1.0f;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz - 1; ++k) {
                float f;
                if (k == piz - 1 || k == piz) f = 1.0f;
                else f = (is_solid_h(ct[I(i, j, k)]) || is_solid_h(ct[I(i, j, k + 1)])) ? 0.0f : 1.0f;
                fz[k + (Nz - 1) * (j + Ny * i)] = f;
            }
}

int alloc_fracs(AcousticFDTDState* st) {
    if (st->d_fx) return FDTD_OK;
    AFDTD_CHECK(dalloc(&st->d_fx, st->N));
    AFDTD_CHECK(dalloc(&st->d_fy, st->N));
    AFDTD_CHECK(dalloc(&st->d_fz, st->N));
    AFDTD_CHECK(hipMemsetAsync(st->d_fx, 0, sizeof(float) * st->N, st->stream));
    AFDTD_CHECK(hipMemsetAsync(st->d_fy, 0, sizeof(float) * st->N, st->stream));
    AFDTD_CHECK(hipMemsetAsync(st->d_fz, 0, sizeof(float) * st->N, st->stream));
    std::vector<float> fx, fy, fz;
    default_fracs(st, fx, fy, fz);
    st->sp.has_closed = 1;
    return upload_fracs(st, fx.data(), fy.data(), fz.data());
}

void destroy_state(AcousticFDTDState* st) {
    if (!st) return;
    if (st->stream) (void)hipStreamSynchronize(st->stream);
    dfree(st->P[0]); dfree(st->P[1]);
    for (int a = 0; a < 3; ++a) { dfree(st->V[a]); dfree(st->V2[a]); }
    dfree(st->stage);
    dfree(st->d_ct); dfree(st->d_fx); dfree(st->d_fy); dfree(st->d_fz);
    dfree(st->d_geo); dfree(st->d_vpl);
    dfree(st->d_g2); dfree(st->d_tw); dfree(st->d_ppre); dfree(st->d_pab);
    dfree(st->d_g2_items); dfree(st->d_g2p);
    if (st->g2_stream) {
        (void)hipStreamSynchronize(st->g2_stream);
        (void)hipStreamDestroy(st->g2_stream);
        (void)hipEventDestroy(st->g2_ev[0]);
        (void)hipEventDestroy(st->g2_ev[1]);
    }
    dfree(st->d_bu); dfree(st->d_kap); dfree(st->d_xi);
    dfree(st->d_alpha); dfree(st->d_pdamp); dfree(st->d_status);
    for (int a = 0; a < 3; ++a) dfree(st->d_w[a]);
    dfree(st->d_ext); dfree(st->d_act); dfree(st->d_above); dfree(st->d_below);
    dfree(st->d_b13); dfree(st->d_vzb); dfree(st->d_vza);
    dfree(st->d_bidx); dfree(st->d_bdrive); dfree(st->d_fidx); dfree(st->d_fval);
    dfree(st->d_bgrp);
    dfree(st->d_pidx); dfree(st->d_psig);
    dfree(st->d_ridx); dfree(st->d_rwgt); dfree(st->d_trace);
    dfree(st->d_gidx); dfree(st->d_gwgt); dfree(st->d_gout);
    dfree(st->d_f2_src_start); dfree(st->d_f2_src); dfree(st->d_f2_tap_start);
    dfree(st->d_f2_tap); dfree(st->d_f2_tapbuf);
    if (st->ev0) (void)hipEventDestroy(st->ev0);
    if (st->ev1) (void)hipEventDestroy(st->ev1);
    if (st->stream) (void)hipStreamDestroy(st->stream);
    delete st;
}

#define TRY(expr)                                                                    \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            afdtd_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));      \
            destroy_state(st);                                                       \
            return nullptr;                                                          \
        }                                                                            \
    } while (0)

AcousticFDTDState* create_impl(int Nx, int Ny, int Nz, float dx, float c, float rho_air,
                               const uint8_t* cell_type, int plate_iz,
                               const uint8_t* plate_active, float plate_mass,
                               float plate_D, float plate_alpha_M, float plate_beta_K,
                               int n_pml, const FdtdOptions* o) {
    // Argument checks of acoustic_fdtd.cpp:341-343.
    if (Nx <= 0 || Ny <= 0 || Nz <= 0 || dx <= 0.0f || c <= 0.0f || cell_type == nullptr ||
        plate_active == nullptr) {
        afdtd_set_error("fdtd_create: invalid arguments");
        return nullptr;
    }
    FdtdOptions def;
    fdtd_default_options(&def);
    if (!o) o = &def;
    if (o->stencil_order != 2 && o->stencil_order != 4) {
        afdtd_set_error("stencil_order must be 2 or 4");
        return nullptr;
    }
    if (o->boundary == FDTD_BOUNDARY_IMPEDANCE && o->stencil_order != 2) {
        afdtd_set_error("impedance boundary requires stencil_order 2");
        return nullptr;
    }
    if (o->storage < 0 || o->storage > 4) {
        afdtd_set_error("unknown storage type");
        return nullptr;
    }
    long long Nll = (long long)Nx * Ny * Nz;
    if (Nll >= (1LL << 31) - (long long)2 * Ny * Nz) {
        afdtd_set_error("grid too large for 32-bit cell indices");
        return nullptr;
    }
    AcousticFDTDState* st = new AcousticFDTDState();   // value-initialised: all zero
    const int N = (int)Nll;
    st->Nx = Nx; st->Ny = Ny; st->Nz = Nz; st->N = N; st->Nplate = Nx * Ny;
    st->Nvx = (Nx - 1) * Ny * Nz; st->Nvy = Nx * (Ny - 1) * Nz; st->Nvz = Nx * Ny * (Nz - 1);
    st->plate_iz = (plate_iz >= 0 && plate_iz < Nz) ? plate_iz : Nz / 2;   // :359
    st->order = o->stencil_order;
    st->boundary = o->boundary;
    st->n_pml = n_pml;
    st->storage = o->storage;
    st->esz = storage_bytes(o->storage);
    st->fbytes = field_bytes(o->storage, Nx, Ny, Nz);
    st->hetero = (o->rho != nullptr) || (o->bulk != nullptr);
    st->dx = dx; st->c = c; st->rho_air = rho_air;

    // dt (acoustic_fdtd.cpp:348) or Courant number over the fastest cell.
    float dt;
    if (o->courant <= 0.0f) {
        dt = 0.77f * dx / (c * 1.732050808f);
    } else {
        double cmax = c;
        if (st->hetero) {
            cmax = 0.0;
            double K0d = (double)((rho_air * c) * c);
            for (int q = 0; q < N; ++q) {
                double r = o->rho ? o->rho[q] : rho_air;
                double kk = o->bulk ? o->bulk[q] : K0d;
                cmax = std::max(cmax, sqrt(kk / r));
            }
        }
        dt = (float)((double)o->courant * (double)dx / cmax);
    }
    st->dt = dt;

    // Cell types: PML retyping (:377-396), plate mask (:413-424).
    st->h_ct.assign(cell_type, cell_type + N);
    st->h_pa.assign(plate_active, plate_active + st->Nplate);
    std::vector<float> alpha_tab(std::max(n_pml, 0) + 1), pdamp_tab(std::max(n_pml, 0) + 1);
    for (int d = 0; d <= std::max(n_pml, 0); ++d) {
        alpha_tab[d] = pml_sigma(d, n_pml, dt);
        float den = 1.0f + alpha_tab[d] * dt;                              // :471
        pdamp_tab[d] = 1.0f / den;                                         // :475
    }
    if (n_pml > 0) {
        for (int i = 0; i < Nx; ++i) {
            int di = pml_depth_h(i, Nx, n_pml);
            for (int j = 0; j < Ny; ++j) {
                int dj = pml_depth_h(j, Ny, n_pml);
                for (int k = 0; k < Nz; ++k) {
                    int dk = pml_depth_h(k, Nz, n_pml);
                    int d = std::max(di, std::max(dj, dk));
                    int q = k + Nz * (j + Ny * i);
                    if (st->h_ct[q] == FDTD_AIR && alpha_tab[d] > 0.0f) st->h_ct[q] = FDTD_PML;
                }
            }
        }
    }
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j) {
            int g = st->plate_iz + Nz * (j + Ny * i);
            if (st->h_pa[j + Ny * i]) {
                if (st->h_ct[g] == FDTD_AIR || st->h_ct[g] == FDTD_PML) st->h_ct[g] = FDTD_PLATE;
            } else if (st->h_ct[g] == FDTD_PLATE) {
                st->h_ct[g] = FDTD_AIR;
            }
        }
    st->has_solid = 0;
    for (int q = 0; q < N; ++q)
        if (is_solid_h(st->h_ct[q])) { st->has_solid = 1; break; }
    int nact = 0;
    for (int p = 0; p < st->Nplate; ++p) nact += plate_active[p] ? 1 : 0;
    st->nact = nact;
    if (nact > 0 && st->storage != FDTD_STORE_F32) {
        afdtd_set_error("plate coupling requires FP32 storage");
        delete st;
        return nullptr;
    }

    // Stencil constants (acoustic_fdtd.cpp:1085-1089, :1176, :889-890).
    StencilParams& sp = st->sp;
    sp.Nx = Nx; sp.Ny = Ny; sp.Nz = Nz;
    sp.order = st->order;
    sp.n_pml = n_pml > 0 ? n_pml : 0;
    sp.impedance = o->boundary == FDTD_BOUNDARY_IMPEDANCE;
    sp.hetero = st->hetero;
    sp.has_solid = st->has_solid;
    sp.has_closed = 0;
    sp.has_frac = 0;
    sp.dt = dt; sp.dx = dx;
    const float inv_dx = 1.0f / dx;
    sp.vel_coeff = dt / (rho_air * dx);
    sp.bulk_coeff = rho_air * c * c * dt * inv_dx;
    sp.C1 = 9.0f / 8.0f;
    sp.C2 = -1.0f / 24.0f;
    for (int f = 0; f < 6; ++f) {
        float R = o->refl[f];
        sp.gside[f] = (1.0f - R) / (1.0f + R);
    }
    sp.rho_air = rho_air;
    sp.K0 = rho_air * c * c;
    sp.dtdx = dt / dx;
    {
        // Homogeneous wall terms (kap/2) * g / sqrt(rho K), one per side.
        const float z0 = sqrtf(rho_air * sp.K0);
        const float hk = 0.5f * sp.bulk_coeff;
        for (int f = 0; f < 6; ++f) sp.xterm[f] = hk * (sp.gside[f] / z0);
    }
    // FP16 stores velocities scaled by 2^round(log2(rho c)) so they sit in
    // the same range as pressures; BF16/FP32 keep the physical value.
    sp.sV = 1.0f;
    if (st->storage == FDTD_STORE_F16) sp.sV = ldexpf(1.0f, (int)lrint(log2((double)rho_air * c)));
    sp.inv_sV = 1.0f / sp.sV;

    TRY(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
    if (resolve_scratch(&st->d_scratch) != FDTD_OK) {
        destroy_state(st);
        return nullptr;
    }
    TRY(hipEventCreate(&st->ev0));
    TRY(hipEventCreate(&st->ev1));
    for (int b = 0; b < 2; ++b) {
        TRY(hipMalloc(&st->P[b], st->fbytes));
        TRY(hipMemsetAsync(st->P[b], 0, st->fbytes, st->stream));
    }
    for (int a = 0; a < 3; ++a) {
        TRY(hipMalloc(&st->V[a], st->fbytes));
        TRY(hipMemsetAsync(st->V[a], 0, st->fbytes, st->stream));
    }
    if (st->storage != FDTD_STORE_F32) TRY(dalloc(&st->stage, N));
    TRY(dalloc(&st->d_status, 2));
    TRY(hipMemsetAsync(st->d_status, 0, 2 * sizeof(int), st->stream));
    TRY(dalloc(&st->d_alpha, alpha_tab.size()));
    TRY(dalloc(&st->d_pdamp, pdamp_tab.size()));
    TRY(hipMemcpyAsync(st->d_alpha, alpha_tab.data(), sizeof(float) * alpha_tab.size(),
                       hipMemcpyHostToDevice, st->stream));
    TRY(hipMemcpyAsync(st->d_pdamp, pdamp_tab.data(), sizeof(float) * pdamp_tab.size(),
                       hipMemcpyHostToDevice, st->stream));
    if (st->has_solid) {
        TRY(dalloc(&st->d_ct, N));
        TRY(hipMemcpyAsync(st->d_ct, st->h_ct.data(), N, hipMemcpyHostToDevice, st->stream));
        TRY(hipStreamSynchronize(st->stream));
        if (alloc_fracs(st) != FDTD_OK) { destroy_state(st); return nullptr; }
    }
    if (st->hetero) {
        // Per-cell buoyancy 1/rho and kappa = K dt/dx; for impedance walls the
        // per-cell wall sum xi = sum_sides (kappa/2) g_side / sqrt(rho K),
        // accumulated in side order x-, x+, y-, y+, z-, z+ (edge cells only).
        std::vector<float> bu(N), kap(N), xi;
        const bool imp = o->boundary == FDTD_BOUNDARY_IMPEDANCE;
        if (imp) xi.assign(N, 0.0f);
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz; ++k) {
                    const int q = k + Nz * (j + Ny * i);
                    const float r = o->rho ? o->rho[q] : rho_air;
                    const float K = o->bulk ? o->bulk[q] : sp.K0;
                    bu[q] = 1.0f / r;
                    kap[q] = K * dt * inv_dx;
                    if (imp && (i == 0 || i == Nx - 1 || j == 0 || j == Ny - 1 || k == 0 || k == Nz - 1)) {
                        const float z = sqrtf(r * K), hk = 0.5f * kap[q];
                        float x = 0.0f;
                        if (i == 0) x = x + hk * (sp.gside[0] / z);
                        if (i == Nx - 1) x = x + hk * (sp.gside[1] / z);
                        if (j == 0) x = x + hk * (sp.gside[2] / z);
                        if (j == Ny - 1) x = x + hk * (sp.gside[3] / z);
                        if (k == 0) x = x + hk * (sp.gside[4] / z);
                        if (k == Nz - 1) x = x + hk * (sp.gside[5] / z);
                        xi[q] = x;
                    }
                }
        TRY(dalloc(&st->d_bu, N)); TRY(dalloc(&st->d_kap, N));
        TRY(hipMemcpyAsync(st->d_bu, bu.data(), sizeof(float) * N, hipMemcpyHostToDevice, st->stream));
        TRY(hipMemcpyAsync(st->d_kap, kap.data(), sizeof(float) * N, hipMemcpyHostToDevice, st->stream));
        if (imp) {
            TRY(dalloc(&st->d_xi, N));
            TRY(hipMemcpyAsync(st->d_xi, xi.data(), sizeof(float) * N, hipMemcpyHostToDevice, st->stream));
        }
        TRY(hipStreamSynchronize(st->stream));
    }

    // Kirchhoff plate (acoustic_fdtd.cpp:434-447, :671-752).
    {
        float m = plate_mass, D = plate_D;
        st->plate_a = (dt * dt) / m;
        float dx4 = dx * dx * dx * dx;
        st->plate_bh = D * (dt * dt) / (m * dx4);
        st->plate_vd = 1.0f - plate_alpha_M * dt;
        st->plate_bk = (dt > 0.0f) ? (plate_beta_K / dt * st->plate_bh) : 0.0f;
        for (int a = 0; a < 3; ++a) {
            TRY(dalloc(&st->d_w[a], st->Nplate));
            TRY(hipMemsetAsync(st->d_w[a], 0, sizeof(float) * st->Nplate, st->stream));
        }
        TRY(dalloc(&st->d_ext, st->Nplate));
        TRY(hipMemsetAsync(st->d_ext, 0, sizeof(float) * st->Nplate, st->stream));
        if (nact > 0) {
            static const int DI[13] = {0, 1, -1, 0, 0, 1, 1, -1, -1, 2, -2, 0, 0};
            static const int DJ[13] = {0, 0, 0, 1, -1, 1, -1, 1, -1, 0, 0, 2, -2};
            std::vector<int> act, above, below, b13, vzb, vza;
            int piz = st->plate_iz;
            for (int i = 0; i < Nx; ++i)
                for (int j = 0; j < Ny; ++j) {
                    int p = j + Ny * i;
                    if (!plate_active[p]) continue;
                    act.push_back(p);
                    above.push_back(piz + 1 < Nz ? (piz + 1) + Nz * p : -1);
                    below.push_back(piz - 1 >= 0 ? (piz - 1) + Nz * p : -1);
                    for (int t = 0; t < 13; ++t) {
                        int ii = i + DI[t], jj = j + DJ[t];
                        if (ii < 0 || ii >= Nx || jj < 0 || jj >= Ny || !plate_active[jj + Ny * ii]) {
                            ii = i; jj = j;
                        }
                        b13.push_back(jj + Ny * ii);
                    }
                    // Vz faces below/above the plate, padded layout (:737-750).
                    vzb.push_back(piz > 0 && !is_solid_h(st->h_ct[(piz - 1) + Nz * p])
                                      ? (piz - 1) + Nz * p : -1);
                    vza.push_back(piz < Nz - 1 && !is_solid_h(st->h_ct[(piz + 1) + Nz * p])
                                      ? piz + Nz * p : -1);
                }
            TRY(dalloc(&st->d_act, nact)); TRY(dalloc(&st->d_above, nact));
            TRY(dalloc(&st->d_below, nact)); TRY(dalloc(&st->d_b13, 13 * nact));
            TRY(dalloc(&st->d_vzb, nact)); TRY(dalloc(&st->d_vza, nact));
            auto up = [&](int* d, const std::vector<int>& h) {
                return hipMemcpyAsync(d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice,
                                      st->stream);
            };
            TRY(up(st->d_act, act)); TRY(up(st->d_above, above)); TRY(up(st->d_below, below));
            TRY(up(st->d_b13, b13)); TRY(up(st->d_vzb, vzb)); TRY(up(st->d_vza, vza));
        }
    }
    TRY(hipStreamSynchronize(st->stream));

    st->fused_tile_y = o->tile_y;
    st->fused_seg_x = o->seg_x;
    st->family = FDTD_KERNEL_GENERIC;
    st->f2_dirty = 1;
    st->kernel_opt = o->kernel;
    if (o->kernel == FDTD_KERNEL_FUSED2 && !fused2_supported(st) && fused4_geo_supported(st)) {
        // walls, fractions or an active plate on the order-4 FP32 grid: the
        // two-step sweep with cell geometry
        st->family = FDTD_KERNEL_FUSED;
        st->fused_geo = 1;
        st->geo_pair = 1;
    } else if (o->kernel == FDTD_KERNEL_FUSED2) {
        if (!fused2_supported(st)) {
            afdtd_set_error("two-step sweep needs no walls, plate cells, face fractions or "
                            "active plate (and for the order-4 stencil FP32 storage and a "
                            "homogeneous medium)");
            destroy_state(st);
            return nullptr;
        }
        st->family = FDTD_KERNEL_FUSED2;
        st->f2_rows = f2_rows_for(o->tile_y);
    } else if (o->kernel == FDTD_KERNEL_AUTO && st->storage == FDTD_STORE_F32 &&
               fused2_supported(st)) {
        // FP32 rooms without walls, fractions or plate, either stencil: two
        // steps per sweep (the single-step sweep is bound by the streaming
        // rate of its byte mix; 16-bit storage halves that mix and keeps the
        // single-step sweep ahead)
        st->family = FDTD_KERNEL_FUSED2;
        st->f2_rows = f2_rows_for(o->tile_y);
    } else if (o->kernel == FDTD_KERNEL_FUSED || o->kernel == FDTD_KERNEL_AUTO) {
        if (fused_supported(st)) {
            st->family = FDTD_KERNEL_FUSED;
        } else if (fused_geo_supported(st)) {
            // walls, plate cells, fractions or an active plate: the sweep with
            // per-cell geometry codes, two steps per pass for the order-4
            // stencil unless the single-step sweep is asked for
            st->family = FDTD_KERNEL_FUSED;
            st->fused_geo = 1;
            st->geo_pair = o->kernel == FDTD_KERNEL_AUTO && fused4_geo_supported(st);
        } else if (o->kernel == FDTD_KERNEL_FUSED) {
            afdtd_set_error("fused kernel does not support this configuration");
            destroy_state(st);
            return nullptr;
        }
    }
    if ((st->storage == FDTD_STORE_BFP16 || st->storage == FDTD_STORE_F16S) &&
        st->family != FDTD_KERNEL_FUSED) {
        afdtd_set_error("block-scaled storage (bfp16, f16s) runs on the single-step fused "
                        "sweep only (no walls, plate cells, face fractions or active plate)");
        destroy_state(st);
        return nullptr;
    }
    if (st->family == FDTD_KERNEL_FUSED || st->family == FDTD_KERNEL_FUSED2) {
        for (int a = 0; a < 3; ++a) {
            TRY(hipMalloc(&st->V2[a], st->fbytes));
            TRY(hipMemsetAsync(st->V2[a], 0, st->fbytes, st->stream));
        }
        if (st->fused_geo) TRY(launch_geo_codes(st));
        if (st->geo_pair) TRY(launch_geo_pair_codes(st));
        TRY(hipStreamSynchronize(st->stream));
    }
    return st;
}

#undef TRY

// par[q] = parity of the number of field-buffer swaps (kernel launches) after
// the first q+1 steps of this call.
int apply_failure(AcousticFDTDState* st, int start_step, int n_steps, int start_cur,
                  float* const w0[3], void* const v0[3], void* const v20[3], int start_trace,
                  int start_phase, const std::vector<uint8_t>& par, int* status_out) {
    int status[2] = {0, 0};
    AFDTD_CHECK(hipMemcpy(status, st->d_status, sizeof(status), hipMemcpyDeviceToHost));
    if (status_out) { status_out[0] = status[0]; status_out[1] = status[1]; }
    if (status[0] == 3 && st->hook_clears_ext && st->nact > 0)
        // the failing step completed its plate update; a hooked run left the
        // accumulator for the next kernel, which no longer runs
        AFDTD_CHECK(launch_fill(st->d_ext, st->Nplate, 0.0f, st->stream));
    if (status[0] == 0) {
        st->step_count = start_step + n_steps;
        return FDTD_OK;
    }
    const int sf = status[1];
    const int completed = status[0] == 3 ? sf - start_step + 1 : sf - start_step;
    st->step_count = start_step + completed;
    const int swaps = completed > 0 ? par[completed - 1] : 0;
    st->cur = start_cur ^ swaps;
    // Plate buffers rotate once per completed step (curr, prev, tmp) <- (tmp, curr, prev).
    float* w[3] = {w0[0], w0[1], w0[2]};
    for (int r = 0; r < completed % 3; ++r) {
        float* t = w[2]; w[2] = w[1]; w[1] = w[0]; w[0] = t;
    }
    for (int a = 0; a < 3; ++a) st->d_w[a] = w[a];
    // The fused sweeps swap the velocity buffers once per launch.
    if (st->family == FDTD_KERNEL_FUSED || st->family == FDTD_KERNEL_FUSED2) {
        for (int a = 0; a < 3; ++a) {
            st->V[a] = swaps ? v20[a] : v0[a];
            st->V2[a] = swaps ? v0[a] : v20[a];
        }
    }
    if (st->trace_cap > 0) st->trace_n = std::min(st->trace_cap, start_trace + completed);
    int checks = 0;
    for (int s = start_step; s < start_step + completed; ++s)
        if (((s + 1) & 0xFF) == 0) ++checks;
    st->phase = (start_phase + checks) % 1013;
    if (status[0] == 3)
        fprintf(stderr, "[fdtd_pressure] UNSTABLE step=%d (sampled stability check)\n",
                st->step_count);
    else if (status[0] != FDTD_STATUS_HOOK)
        fprintf(stderr, "[fdtd_plate] UNSTABLE step=%d (plate displacement guard)\n",
                st->step_count);
    fflush(stderr);
    AFDTD_CHECK(hipMemsetAsync(st->d_status, 0, 2 * sizeof(int), st->stream));
    AFDTD_CHECK(hipStreamSynchronize(st->stream));
    return FDTD_ERR_UNSTABLE;
}

// Enqueue one full step with global index s.  keep_ext: the plate force
// accumulator is left for the caller's next hooked kernel to clear.
int enqueue_step(AcousticFDTDState* st, int s, hipEvent_t ea, hipEvent_t eb,
                 bool keep_ext = false) {
    if (st->nact > 0) {
        AFDTD_CHECK(launch_plate_step(st, s));
        if (!keep_ext) AFDTD_CHECK(launch_fill(st->d_ext, st->Nplate, 0.0f, st->stream));
        float* t = st->d_w[2]; st->d_w[2] = st->d_w[1]; st->d_w[1] = st->d_w[0]; st->d_w[0] = t;
    }
    if (ea) AFDTD_CHECK(hipEventRecord(ea, st->stream));
    if (st->family == FDTD_KERNEL_FUSED || st->family == FDTD_KERNEL_FUSED2) {
        st->side_step = s;
        AFDTD_CHECK(launch_fused_step(st));
    } else {
        AFDTD_CHECK(launch_velocity_generic(st));
        AFDTD_CHECK(launch_plate_inject(st));
        AFDTD_CHECK(launch_pressure_generic(st));
    }
    if (eb) AFDTD_CHECK(hipEventRecord(eb, st->stream));
    st->cur ^= 1;
    int row = -1;
    if (st->d_trace && st->trace_n < st->trace_cap) row = st->trace_n++;
    AFDTD_CHECK(launch_post_step(st, s, row));
    if (((s + 1) & 0xFF) == 0) {
        AFDTD_CHECK(launch_stability_check(st, st->phase, s));
        st->phase = (st->phase + 1) % 1013;
    }
    return FDTD_OK;
}

// Enqueue steps s and s+1 as one two-step sweep (step s+1 is not a
// stability-check step: the caller schedules pairs so).
int enqueue_pair(AcousticFDTDState* st, int s, hipEvent_t ea, hipEvent_t eb) {
    if (ea) AFDTD_CHECK(hipEventRecord(ea, st->stream));
    AFDTD_CHECK(launch_fused2_pair(st, s));
    if (eb) AFDTD_CHECK(hipEventRecord(eb, st->stream));
    st->cur ^= 1;
    int row0 = -1, row1 = -1;
    if (st->d_trace && st->trace_n < st->trace_cap) row0 = st->trace_n++;
    if (st->d_trace && st->trace_n < st->trace_cap) row1 = st->trace_n++;
    AFDTD_CHECK(launch_fused2_taps(st, row0));
    AFDTD_CHECK(launch_post_step(st, s + 1, row1));
    if (((s + 2) & 0xFF) == 0) {
        AFDTD_CHECK(launch_stability_check(st, st->phase, s + 1));
        st->phase = (st->phase + 1) % 1013;
    }
    return FDTD_OK;
}

// Enqueue steps s and s+1 as one two-step sweep with cell geometry (the
// caller schedules pairs away from stability-check steps, hooks and side
// jobs).  Step s's plate kernel runs first, as in enqueue_step; step s+1's
// plate update runs inside the pair launch.
int enqueue_geo_pair(AcousticFDTDState* st, int s, hipEvent_t ea, hipEvent_t eb) {
    if (st->nact > 0) {
        AFDTD_CHECK(launch_plate_step(st, s));
        AFDTD_CHECK(launch_fill(st->d_ext, st->Nplate, 0.0f, st->stream));
        float* t = st->d_w[2]; st->d_w[2] = st->d_w[1]; st->d_w[1] = st->d_w[0]; st->d_w[0] = t;
    }
    int row0 = -1, row1 = -1;
    if (st->d_trace && st->trace_n < st->trace_cap) row0 = st->trace_n++;
    if (st->d_trace && st->trace_n < st->trace_cap) row1 = st->trace_n++;
    if (ea) AFDTD_CHECK(hipEventRecord(ea, st->stream));
    AFDTD_CHECK(launch_fused4_geo_pair(st, s, row0));
    if (eb) AFDTD_CHECK(hipEventRecord(eb, st->stream));
    st->cur ^= 1;
    AFDTD_CHECK(launch_post_step(st, s + 1, row1));
    if (((s + 2) & 0xFF) == 0) {
        AFDTD_CHECK(launch_stability_check(st, st->phase, s + 1));
        st->phase = (st->phase + 1) % 1013;
    }
    return FDTD_OK;
}

int step_impl(AcousticFDTDState* st, int n_steps, float* ms_total, float* ms_stencil,
              const StepHook* hook = nullptr, int* status_out = nullptr) {
    const int start_step = st->step_count, start_cur = st->cur;
    const int start_trace = st->trace_n, start_phase = st->phase;
    float* w0[3] = {st->d_w[0], st->d_w[1], st->d_w[2]};
    void* v0[3] = {st->V[0], st->V[1], st->V[2]};
    void* v20[3] = {st->V2[0], st->V2[1], st->V2[2]};
    std::vector<hipEvent_t> evs;
    if (ms_stencil) {
        evs.resize(2 * (size_t)n_steps);
        for (auto& e : evs) AFDTD_CHECK(hipEventCreate(&e));
    }
    if (ms_total) AFDTD_CHECK(hipEventRecord(st->ev0, st->stream));
    std::vector<uint8_t> par((size_t)n_steps, 0);
    int launches = 0;
    if (st->family == FDTD_KERNEL_FUSED2 || (st->geo_pair && !hook)) {
        int rc = fused2_prepare(st);
        if (rc) return rc;
    }
    for (int q = 0; q < n_steps;) {
        const int s = start_step + q;
        hipEvent_t ea = ms_stencil ? evs[2 * launches] : nullptr;
        hipEvent_t eb = ms_stencil ? evs[2 * launches + 1] : nullptr;
        if (hook) {
            int rc = (*hook)(q, s);
            if (rc) return rc;
        }
        const bool pairable = !hook && st->f2_ok && q + 1 < n_steps && ((s + 1) & 0xFF) != 0;
        const bool pair = pairable && st->family == FDTD_KERNEL_FUSED2 && st->nact == 0;
        const bool gpair = pairable && st->geo_pair && st->side_n == 0;
        const bool keep_ext = hook && st->hook_clears_ext && q + 1 < n_steps;
        int rc = gpair ? enqueue_geo_pair(st, s, ea, eb)
                 : pair ? enqueue_pair(st, s, ea, eb)
                        : enqueue_step(st, s, ea, eb, keep_ext);
        if (rc) return rc;
        ++launches;
        const int adv = (pair || gpair) ? 2 : 1;
        for (int u = 0; u < adv; ++u) par[q + u] = (uint8_t)(launches & 1);
        q += adv;
    }
    if (hook) {
        int rc = (*hook)(n_steps, start_step + n_steps);
        if (rc) return rc;
    }
    if (ms_total) AFDTD_CHECK(hipEventRecord(st->ev1, st->stream));
    AFDTD_CHECK(hipStreamSynchronize(st->stream));
    if (ms_total) AFDTD_CHECK(hipEventElapsedTime(ms_total, st->ev0, st->ev1));
    if (ms_stencil) {
        // average stencil time per step (a two-step launch counts for two)
        double sum = 0.0;
        for (int q = 0; q < launches; ++q) {
            float t = 0.0f;
            AFDTD_CHECK(hipEventElapsedTime(&t, evs[2 * q], evs[2 * q + 1]));
            sum += t;
        }
        *ms_stencil = (float)(sum / n_steps);
        for (auto& e : evs) (void)hipEventDestroy(e);
    }
    return apply_failure(st, start_step, n_steps, start_cur, w0, v0, v20, start_trace,
                         start_phase, par, status_out);
}

}  // namespace

int afdtd_step_hooked(AcousticFDTDState* st, int n_steps, const StepHook& hook, float* ms_total,
                      int* status_out) {
    if (!st) return FDTD_ERR_NULL;
    if (status_out) { status_out[0] = 0; status_out[1] = 0; }
    if (n_steps <= 0) return FDTD_OK;
    return step_impl(st, n_steps, ms_total, nullptr, &hook, status_out);
}

extern "C" {

SK_API void fdtd_default_options(FdtdOptions* o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->struct_size = (int)sizeof(FdtdOptions);
    o->stencil_order = FDTD_ORDER_4;
    o->courant = 0.0f;
    o->boundary = FDTD_BOUNDARY_RIGID_PML;
    for (int f = 0; f < 6; ++f) o->refl[f] = 1.0f;
    o->storage = FDTD_STORE_F32;
    o->kernel = FDTD_KERNEL_AUTO;
}

SK_API AcousticFDTDState* fdtd_create(int Nx, int Ny, int Nz, float dx, float c, float rho_air,
                                      const uint8_t* cell_type, int plate_iz,
                                      const uint8_t* plate_active, float plate_mass_density,
                                      float plate_stiffness_D, float plate_alpha_M,
                                      float plate_beta_K, int n_pml) {
    return create_impl(Nx, Ny, Nz, dx, c, rho_air, cell_type, plate_iz, plate_active,
                       plate_mass_density, plate_stiffness_D, plate_alpha_M, plate_beta_K,
                       n_pml, nullptr);
}

SK_API AcousticFDTDState* fdtd_create_ex(int Nx, int Ny, int Nz, float dx, float c,
                                         float rho_air, const uint8_t* cell_type, int plate_iz,
                                         const uint8_t* plate_active, float plate_mass_density,
                                         float plate_stiffness_D, float plate_alpha_M,
                                         float plate_beta_K, int n_pml, const FdtdOptions* opt) {
    if (opt && opt->struct_size != (int)sizeof(FdtdOptions)) {
        afdtd_set_error("FdtdOptions.struct_size mismatch");
        return nullptr;
    }
    return create_impl(Nx, Ny, Nz, dx, c, rho_air, cell_type, plate_iz, plate_active,
                       plate_mass_density, plate_stiffness_D, plate_alpha_M, plate_beta_K,
                       n_pml, opt);
}

SK_API void fdtd_destroy(AcousticFDTDState* st) { destroy_state(st); }

SK_API int fdtd_set_bridge_sources(AcousticFDTDState* st, int n, const int* idx,
                                   const float* wgt) {
    if (!st || n <= 0 || !idx || !wgt) return FDTD_ERR_NULL;
    float wsum = 0.0f;                                                  // :815-817
    for (int i = 0; i < n; ++i) wsum += fabsf(wgt[i]);
    float wscale = (wsum > 1e-12f) ? 1.0f / wsum : 1.0f;
    st->h_bidx.assign(idx, idx + n);
    st->h_bwgt.resize(n);
    for (int i = 0; i < n; ++i) st->h_bwgt[i] = wgt[i] * wscale;
    dfree(st->d_bidx); dfree(st->d_bdrive); dfree(st->d_bgrp);
    st->d_bidx = nullptr; st->d_bdrive = nullptr; st->d_bgrp = nullptr;
    st->nbsrc = 0;
    std::vector<int> grp;
    const int ng = force_groups(idx, n, st->Nz, nullptr, st->Nplate, grp);
    AFDTD_CHECK(dalloc(&st->d_bidx, n));
    AFDTD_CHECK(dalloc(&st->d_bdrive, n));
    AFDTD_CHECK(dalloc(&st->d_bgrp, grp.size()));
    AFDTD_CHECK(hipMemcpy(st->d_bidx, idx, sizeof(int) * n, hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(st->d_bgrp, grp.data(), sizeof(int) * grp.size(), hipMemcpyHostToDevice));
    st->nbgrp = ng;
    st->nbsrc = n;
    return FDTD_OK;
}

SK_API int fdtd_inject_bridge(AcousticFDTDState* st, float a, float b, float c) {
    (void)st; (void)a; (void)b; (void)c;
    return FDTD_ERR_UNSTABLE;                                           // :874-885
}

SK_API int fdtd_inject_bridge_drive(AcousticFDTDState* st, const float* drives, float fs) {
    if (!st) return FDTD_ERR_NULL;
    if (st->nbsrc == 0) return FDTD_OK;                                 // :1705
    if (!drives) return FDTD_ERR_NULL;
    AFDTD_CHECK(hipMemcpyAsync(st->d_bdrive, drives, sizeof(float) * st->nbsrc,
                               hipMemcpyHostToDevice, st->stream));
    float inv_dx2 = 1.0f / (st->dx * st->dx);
    AFDTD_CHECK(launch_accumulate_groups(st->d_ext, st->d_bgrp, st->nbgrp, st->d_bdrive, fs,
                                         inv_dx2, st->d_status, st->stream));
    return sync(st);
}

SK_API int fdtd_inject_plate_force(AcousticFDTDState* st, int n, const int* pidx,
                                   const float* wgt, float total) {
    if (!st || !pidx || !wgt) return FDTD_ERR_NULL;
    if (n <= 0 || total == 0.0f) return FDTD_OK;
    std::vector<int> grp;                                               // :1735-1740
    const int ng = force_groups(pidx, n, 1, st->h_pa.data(), st->Nplate, grp);
    if (ng == 0) return FDTD_OK;
    if ((int)grp.size() > st->fcap || n > st->fcap) {
        dfree(st->d_fidx); dfree(st->d_fval);
        st->d_fidx = nullptr; st->d_fval = nullptr; st->fcap = 0;
        const int cap = std::max((int)grp.size(), n);
        AFDTD_CHECK(dalloc(&st->d_fidx, cap));
        AFDTD_CHECK(dalloc(&st->d_fval, cap));
        st->fcap = cap;
    }
    AFDTD_CHECK(hipMemcpyAsync(st->d_fidx, grp.data(), sizeof(int) * grp.size(),
                               hipMemcpyHostToDevice, st->stream));
    AFDTD_CHECK(hipMemcpyAsync(st->d_fval, wgt, sizeof(float) * n, hipMemcpyHostToDevice, st->stream));
    float inv_dx2 = 1.0f / (st->dx * st->dx);
    AFDTD_CHECK(launch_accumulate_groups(st->d_ext, st->d_fidx, ng, st->d_fval, total, inv_dx2,
                                         st->d_status, st->stream));
    return sync(st);
}

SK_API int fdtd_set_face_fractions(AcousticFDTDState* st, const float* fx, const float* fy,
                                   const float* fz) {
    if (!st) return FDTD_ERR_NULL;
    if ((st->storage == FDTD_STORE_BFP16 || st->storage == FDTD_STORE_F16S) && (fx || fy || fz)) {
        afdtd_set_error("face fractions need FP32, BF16 or FP16 storage (no block-scaled "
                        "storage: bfp16, f16s)");
        return FDTD_ERR_ARG;
    }
    int rc = alloc_fracs(st);
    if (rc) return rc;
    rc = upload_fracs(st, fx, fy, fz);
    if (rc) return rc;
    if (fx || fy || fz) {
        st->has_frac = 1;
        st->sp.has_frac = 1;
    }
    // The plain sweeps carry no geometry: fractions move a fused handle to
    // the single-step sweep with geometry codes (FP32), else to the generic
    // kernels.  Both fused families keep the same double-buffered velocities.
    if ((st->family == FDTD_KERNEL_FUSED || st->family == FDTD_KERNEL_FUSED2) &&
        !st->fused_geo && !fused_supported(st)) {
        if (fused_geo_supported(st)) {
            st->family = FDTD_KERNEL_FUSED;
            st->fused_geo = 1;
            st->geo_pair = st->kernel_opt != FDTD_KERNEL_FUSED && fused4_geo_supported(st);
        } else {
            st->family = FDTD_KERNEL_GENERIC;
        }
    }
    if (st->fused_geo) {
        AFDTD_CHECK(launch_geo_codes(st));
        if (st->geo_pair) AFDTD_CHECK(launch_geo_pair_codes(st));
        return sync(st);
    }
    return FDTD_OK;
}

SK_API int fdtd_sample_plate_weighted(const AcousticFDTDState* cst, int n, const int* pidx,
                                      const float* wgt, float* out) {
    if (!cst || !pidx || !wgt || !out) return FDTD_ERR_NULL;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    std::vector<float> w(st->Nplate);
    AFDTD_CHECK(hipMemcpyAsync(w.data(), st->d_w[0], sizeof(float) * st->Nplate,
                               hipMemcpyDeviceToHost, st->stream));
    int rc = sync(st);
    if (rc) return rc;
    float acc = 0.0f;                                                   // :1752-1757
    for (int i = 0; i < n; ++i) {
        int p = pidx[i];
        if (p < 0 || p >= st->Nplate) continue;
        acc += w[p] * wgt[i];
    }
    *out = acc;
    return FDTD_OK;
}

SK_API int fdtd_precompute_pressure_samplers(const AcousticFDTDState* st, int n,
                                             const float* xyz, int* idx8, float* wgt8) {
    if (!st || !xyz || !idx8 || !wgt8) return FDTD_ERR_NULL;
    for (int r = 0; r < n; ++r) p_sampler(st->Nx, st->Ny, st->Nz, xyz + 3 * r, idx8 + 8 * r, wgt8 + 8 * r);
    return FDTD_OK;
}

SK_API int fdtd_precompute_velocity_samplers(const AcousticFDTDState* st, int n,
                                             const float* xyz, int* xi, float* xw, int* yi,
                                             float* yw, int* zi, float* zw) {
    if (!st || !xyz) return FDTD_ERR_NULL;
    if (!xi || !xw || !yi || !yw || !zi || !zw) return FDTD_ERR_NULL;
    int* I[3] = {xi, yi, zi};
    float* W[3] = {xw, yw, zw};
    for (int ax = 0; ax < 3; ++ax)
        for (int r = 0; r < n; ++r)
            v_sampler(st->Nx, st->Ny, st->Nz, ax, xyz + 3 * r, I[ax] + 8 * r, W[ax] + 8 * r);
    return FDTD_OK;
}

SK_API int fdtd_sample_pressure_precomputed(const AcousticFDTDState* cst, int n, const int* idx8,
                                            const float* wgt8, float* out) {
    if (!cst || !idx8 || !wgt8 || !out) return FDTD_ERR_NULL;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    return gather_host(st, 0, idx8, wgt8, n, 8, 0, out);
}

SK_API int fdtd_sample_pressure_precomputed_noalloc(const AcousticFDTDState* cst, int n,
                                                    const int* idx8, const float* wgt8,
                                                    float* out) {
    if (!cst || !idx8 || !wgt8 || !out) return FDTD_ERR_NULL;
    if (n > FDTD_NOALLOC_MAX_RECEIVERS) {
        afdtd_set_error("fdtd_sample_pressure_precomputed_noalloc: " + std::to_string(n) +
                        " receivers exceed the allocation-free maximum of " +
                        std::to_string(FDTD_NOALLOC_MAX_RECEIVERS));
        return FDTD_ERR_DIM;
    }
    if (n <= 0) return FDTD_OK;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    if (!st->d_scratch) {
        afdtd_set_error("fdtd_sample_pressure_precomputed_noalloc: scratch not resolved");
        return FDTD_ERR_DEVICE;
    }
    int rc = check_indices(st, 0, idx8, (size_t)8 * n);
    if (rc) return rc;
    int* d_idx = (int*)(st->d_scratch + kScrIdx);
    float* d_wgt = (float*)(st->d_scratch + kScrWgt);
    float* d_out = (float*)(st->d_scratch + kScrOut);
    // the stencils and outputs overwrite [0, 8n) and [0, n): nothing of the
    // previous call's contents is read
    AFDTD_CHECK(hipMemcpyAsync(d_idx, idx8, sizeof(int) * 8 * n, hipMemcpyHostToDevice,
                               st->stream));
    AFDTD_CHECK(hipMemcpyAsync(d_wgt, wgt8, sizeof(float) * 8 * n, hipMemcpyHostToDevice,
                               st->stream));
    AFDTD_CHECK(launch_gather(st->P[st->cur], st->storage, 1.0f, d_idx, d_wgt, n, 8, 0, d_out,
                              st->stream, st->N, st->Nz));
    AFDTD_CHECK(hipMemcpyAsync(out, d_out, sizeof(float) * n, hipMemcpyDeviceToHost, st->stream));
    return sync(st);
}

SK_API int fdtd_sample_velocity_precomputed(const AcousticFDTDState* cst, int n,
                                            const int* xi, const float* xw, const int* yi,
                                            const float* yw, const int* zi, const float* zw,
                                            float* ox, float* oy, float* oz) {
    if (!cst || !xi || !xw || !yi || !yw || !zi || !zw) return FDTD_ERR_NULL;
    if (!ox || !oy || !oz) return FDTD_ERR_NULL;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    const int* I[3] = {xi, yi, zi};
    const float* W[3] = {xw, yw, zw};
    float* O[3] = {ox, oy, oz};
    std::vector<int> pad((size_t)8 * std::max(n, 0));
    for (int ax = 0; ax < 3; ++ax) {
        for (int q = 0; q < 8 * n; ++q)
            pad[q] = packed_to_padded(ax, I[ax][q], st->Nx, st->Ny, st->Nz);
        int rc = gather_host(st, 1 + ax, pad.data(), W[ax], n, 8, 1, O[ax]);
        if (rc) return rc;
    }
    return FDTD_OK;
}

SK_API int fdtd_step(AcousticFDTDState* st, int n_steps) {
    if (!st) return FDTD_ERR_NULL;
    if (n_steps <= 0) return FDTD_OK;
    return step_impl(st, n_steps, nullptr, nullptr);
}

SK_API int fdtd_step_timed(AcousticFDTDState* st, int n_steps, float* ms_total,
                           float* ms_stencil_avg) {
    if (!st || !ms_total || !ms_stencil_avg) return FDTD_ERR_NULL;
    if (n_steps <= 0) return FDTD_OK;
    return step_impl(st, n_steps, ms_total, ms_stencil_avg);
}

SK_API int fdtd_get_step_count(const AcousticFDTDState* st) { return st ? st->step_count : 0; }
SK_API float fdtd_get_dt(const AcousticFDTDState* st) { return st ? st->dt : 0.0f; }
SK_API int fdtd_get_kernel_family(const AcousticFDTDState* st) { return st ? st->family : 0; }

SK_API int fdtd_get_sweep_steps(const AcousticFDTDState* st) {
    if (!st) return 0;
    return (st->family == FDTD_KERNEL_FUSED2 || st->geo_pair) ? 2 : 1;
}

SK_API int fdtd_reset(AcousticFDTDState* st) {
    if (!st) return FDTD_ERR_NULL;                                      // :1297-1310
    for (int b = 0; b < 2; ++b)
        AFDTD_CHECK(hipMemsetAsync(st->P[b], 0, st->fbytes, st->stream));
    for (int a = 0; a < 3; ++a)
        AFDTD_CHECK(hipMemsetAsync(st->V[a], 0, st->fbytes, st->stream));
    AFDTD_CHECK(hipMemsetAsync(st->d_w[0], 0, sizeof(float) * st->Nplate, st->stream));
    AFDTD_CHECK(hipMemsetAsync(st->d_w[1], 0, sizeof(float) * st->Nplate, st->stream));
    AFDTD_CHECK(hipMemsetAsync(st->d_ext, 0, sizeof(float) * st->Nplate, st->stream));
    st->step_count = 0;
    st->trace_n = 0;
    return sync(st);
}

SK_API int fdtd_get_pressure_field(const AcousticFDTDState* cst, float* out, int len) {
    if (!cst || !out) return FDTD_ERR_NULL;
    if (len != cst->N) return FDTD_ERR_DIM;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    const float* src = field_f32(st, 0);
    AFDTD_CHECK(hipMemcpyAsync(out, src, sizeof(float) * st->N, hipMemcpyDeviceToHost, st->stream));
    return sync(st);
}

SK_API int fdtd_get_plate_displacement(const AcousticFDTDState* cst, float* out, int len) {
    if (!cst || !out) return FDTD_ERR_NULL;
    if (len != cst->Nplate) return FDTD_ERR_DIM;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    AFDTD_CHECK(hipMemcpyAsync(out, st->d_w[0], sizeof(float) * st->Nplate, hipMemcpyDeviceToHost,
                               st->stream));
    return sync(st);
}

SK_API int fdtd_sample_plate_displacement_batch(const AcousticFDTDState* cst, int n,
                                                const int* idx4, const float* wgt4, float* out) {
    if (!cst || !idx4 || !wgt4 || !out) return FDTD_ERR_NULL;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    std::vector<float> w(st->Nplate);
    AFDTD_CHECK(hipMemcpyAsync(w.data(), st->d_w[0], sizeof(float) * st->Nplate,
                               hipMemcpyDeviceToHost, st->stream));
    int rc = sync(st);
    if (rc) return rc;
    for (int s = 0; s < n; ++s) {                                       // :1342-1351
        const int* id = idx4 + 4 * s;
        const float* g = wgt4 + 4 * s;
        out[s] = g[0] * w[id[0]] + g[1] * w[id[1]] + g[2] * w[id[2]] + g[3] * w[id[3]];
    }
    return FDTD_OK;
}

SK_API int fdtd_sample_pressure(const AcousticFDTDState* cst, int n, const float* xyz, float* out) {
    if (!cst || !xyz || !out) return FDTD_ERR_NULL;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    if (n <= 0) return FDTD_OK;
    std::vector<int> idx((size_t)8 * n);
    std::vector<float> wgt((size_t)8 * n);
    for (int r = 0; r < n; ++r) p_sampler(st->Nx, st->Ny, st->Nz, xyz + 3 * r, &idx[8 * r], &wgt[8 * r]);
    return gather_host(st, 0, idx.data(), wgt.data(), n, 8, 1, out);
}

SK_API void fdtd_get_velocity_dims(const AcousticFDTDState* st, int* a, int* b, int* c) {
    if (!st) { *a = *b = *c = 0; return; }
    *a = st->Nvx; *b = st->Nvy; *c = st->Nvz;
}

static int get_velocity(AcousticFDTDState* st, int ax, float* out, int len) {
    int Nx = st->Nx, Ny = st->Ny, Nz = st->Nz;
    int need = ax == 0 ? st->Nvx : (ax == 1 ? st->Nvy : st->Nvz);
    if (len != need) return FDTD_ERR_DIM;
    if (need <= 0) return FDTD_OK;
    const float* src = field_f32(st, 1 + ax);
    if (ax == 0) {
        AFDTD_CHECK(hipMemcpyAsync(out, src, sizeof(float) * need, hipMemcpyDeviceToHost, st->stream));
    } else if (ax == 1) {
        AFDTD_CHECK(hipMemcpy2DAsync(out, sizeof(float) * (Ny - 1) * Nz, src, sizeof(float) * Ny * Nz,
                                     sizeof(float) * (Ny - 1) * Nz, Nx, hipMemcpyDeviceToHost,
                                     st->stream));
    } else {
        AFDTD_CHECK(hipMemcpy2DAsync(out, sizeof(float) * (Nz - 1), src, sizeof(float) * Nz,
                                     sizeof(float) * (Nz - 1), (size_t)Nx * Ny,
                                     hipMemcpyDeviceToHost, st->stream));
    }
    return sync(st);
}

SK_API int fdtd_get_velocity_x(const AcousticFDTDState* st, float* out, int len) {
    if (!st || !out) return FDTD_ERR_NULL;
    return get_velocity(const_cast<AcousticFDTDState*>(st), 0, out, len);
}
SK_API int fdtd_get_velocity_y(const AcousticFDTDState* st, float* out, int len) {
    if (!st || !out) return FDTD_ERR_NULL;
    return get_velocity(const_cast<AcousticFDTDState*>(st), 1, out, len);
}
SK_API int fdtd_get_velocity_z(const AcousticFDTDState* st, float* out, int len) {
    if (!st || !out) return FDTD_ERR_NULL;
    return get_velocity(const_cast<AcousticFDTDState*>(st), 2, out, len);
}

SK_API int fdtd_sample_velocity(const AcousticFDTDState* cst, int n, const float* xyz, float* ox,
                                float* oy, float* oz) {
    if (!cst || !xyz || !ox || !oy || !oz) return FDTD_ERR_NULL;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    if (n <= 0) return FDTD_OK;
    std::vector<int> idx((size_t)8 * n);
    std::vector<float> wgt((size_t)8 * n);
    float* O[3] = {ox, oy, oz};
    for (int ax = 0; ax < 3; ++ax) {
        for (int r = 0; r < n; ++r) {
            v_sampler(st->Nx, st->Ny, st->Nz, ax, xyz + 3 * r, &idx[8 * r], &wgt[8 * r]);
            for (int t = 0; t < 8; ++t)
                idx[8 * r + t] = packed_to_padded(ax, idx[8 * r + t], st->Nx, st->Ny, st->Nz);
        }
        int rc = gather_host(st, 1 + ax, idx.data(), wgt.data(), n, 8, 1, O[ax]);
        if (rc) return rc;
    }
    return FDTD_OK;
}

SK_API int fdtd_get_surface_emission(const AcousticFDTDState* st, int n, const float* xyz,
                                     const float* nrm, float* oP, float* ovn) {
    if (!st || !xyz || !nrm || !oP || !ovn) return FDTD_ERR_NULL;
    if (n <= 0) return FDTD_OK;
    int rc = fdtd_sample_pressure(st, n, xyz, oP);                      // :1546-1561
    if (rc) return rc;
    std::vector<float> vx(n), vy(n), vz(n);
    rc = fdtd_sample_velocity(st, n, xyz, vx.data(), vy.data(), vz.data());
    if (rc) return rc;
    for (int s = 0; s < n; ++s)
        ovn[s] = vx[s] * nrm[3 * s] + vy[s] * nrm[3 * s + 1] + vz[s] * nrm[3 * s + 2];
    return FDTD_OK;
}

// ── extensions ────────────────────────────────────────────────────────────

SK_API int fdtd_set_point_sources(AcousticFDTDState* st, int n, const int* cell_idx,
                                  const float* signals, int n_samples) {
    if (!st) return FDTD_ERR_NULL;
    dfree(st->d_pidx); dfree(st->d_psig);
    st->d_pidx = nullptr; st->d_psig = nullptr; st->npsrc = 0; st->plen = 0;
    st->h_pidx.clear();
    st->f2_dirty = 1;
    if (n <= 0) return FDTD_OK;
    if (!cell_idx || !signals || n_samples <= 0) return FDTD_ERR_NULL;
    for (int s = 0; s < n; ++s) {
        if (cell_idx[s] < 0 || cell_idx[s] >= st->N || is_solid_h(st->h_ct[cell_idx[s]])) {
            afdtd_set_error("point source outside the grid or on a solid cell");
            return FDTD_ERR_DIM;
        }
    }
    AFDTD_CHECK(dalloc(&st->d_pidx, n));
    AFDTD_CHECK(dalloc(&st->d_psig, (size_t)n * n_samples));
    AFDTD_CHECK(hipMemcpy(st->d_pidx, cell_idx, sizeof(int) * n, hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(st->d_psig, signals, sizeof(float) * n * (size_t)n_samples,
                          hipMemcpyHostToDevice));
    st->npsrc = n;
    st->plen = n_samples;
    st->h_pidx.assign(cell_idx, cell_idx + n);
    return FDTD_OK;
}

SK_API int fdtd_set_receivers(AcousticFDTDState* st, int n, const float* xyz) {
    if (!st) return FDTD_ERR_NULL;
    dfree(st->d_ridx); dfree(st->d_rwgt); dfree(st->d_trace);
    st->d_ridx = nullptr; st->d_rwgt = nullptr; st->d_trace = nullptr;
    st->nrec = 0; st->trace_cap = 0; st->trace_n = 0;
    st->h_ridx.clear();
    st->f2_dirty = 1;
    if (n <= 0) return FDTD_OK;
    if (!xyz) return FDTD_ERR_NULL;
    std::vector<int> idx((size_t)8 * n);
    std::vector<float> wgt((size_t)8 * n);
    for (int r = 0; r < n; ++r) p_sampler(st->Nx, st->Ny, st->Nz, xyz + 3 * r, &idx[8 * r], &wgt[8 * r]);
    AFDTD_CHECK(dalloc(&st->d_ridx, (size_t)8 * n));
    AFDTD_CHECK(dalloc(&st->d_rwgt, (size_t)8 * n));
    AFDTD_CHECK(hipMemcpy(st->d_ridx, idx.data(), sizeof(int) * 8 * n, hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(st->d_rwgt, wgt.data(), sizeof(float) * 8 * n, hipMemcpyHostToDevice));
    st->nrec = n;
    st->h_ridx = idx;
    return FDTD_OK;
}

SK_API int fdtd_record_receivers(AcousticFDTDState* st, int max_steps) {
    if (!st) return FDTD_ERR_NULL;
    if (st->nrec <= 0 || max_steps <= 0) return FDTD_ERR_DIM;
    dfree(st->d_trace);
    st->d_trace = nullptr;
    AFDTD_CHECK(dalloc(&st->d_trace, (size_t)max_steps * st->nrec));
    st->trace_cap = max_steps;
    st->trace_n = 0;
    return FDTD_OK;
}

SK_API int fdtd_get_receiver_traces(const AcousticFDTDState* cst, float* out, int len,
                                    int* n_recorded) {
    if (!cst || !out) return FDTD_ERR_NULL;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    size_t need = (size_t)st->trace_n * st->nrec;
    if ((size_t)len < need) return FDTD_ERR_DIM;
    if (need)
        AFDTD_CHECK(hipMemcpyAsync(out, st->d_trace, sizeof(float) * need, hipMemcpyDeviceToHost,
                                   st->stream));
    if (n_recorded) *n_recorded = st->trace_n;
    return sync(st);
}

SK_API int fdtd_energy(const AcousticFDTDState* cst, double* out2) {
    if (!cst || !out2) return FDTD_ERR_NULL;
    AcousticFDTDState* st = const_cast<AcousticFDTDState*>(cst);
    double* d = nullptr;
    AFDTD_CHECK(dalloc(&d, 2));
    AFDTD_CHECK(hipMemsetAsync(d, 0, 2 * sizeof(double), st->stream));
    AFDTD_CHECK(launch_energy(st, d));
    AFDTD_CHECK(hipMemcpyAsync(out2, d, 2 * sizeof(double), hipMemcpyDeviceToHost, st->stream));
    int rc = sync(st);
    dfree(d);
    return rc;
}

SK_API int fdtd_set_state(AcousticFDTDState* st, const float* P, const float* vx, const float* vy,
                          const float* vz) {
    if (!st) return FDTD_ERR_NULL;
    int Nx = st->Nx, Ny = st->Ny, Nz = st->Nz;
    float* tmp = nullptr;
    AFDTD_CHECK(dalloc(&tmp, st->N));
    auto put = [&](void* dst, float scale) -> int {
        AFDTD_CHECK(launch_from_f32(tmp, dst, st->storage, scale, (size_t)st->N, st->stream,
                                    st->Nz));
        return FDTD_OK;
    };
    int rc = FDTD_OK;
    if (P) {
        std::vector<float> h(P, P + st->N);
        for (int q = 0; q < st->N; ++q)
            if (is_solid_h(st->h_ct[q])) h[q] = 0.0f;   // solid cells hold P = 0
        AFDTD_CHECK(hipMemcpyAsync(tmp, h.data(), sizeof(float) * st->N, hipMemcpyHostToDevice, st->stream));
        rc = rc ? rc : put(st->P[st->cur], 1.0f);
        AFDTD_CHECK(hipStreamSynchronize(st->stream));
    }
    const float* src[3] = {vx, vy, vz};
    for (int ax = 0; ax < 3 && !rc; ++ax) {
        if (!src[ax]) continue;
        AFDTD_CHECK(hipMemsetAsync(tmp, 0, sizeof(float) * st->N, st->stream));
        if (ax == 0 && st->Nvx > 0)
            AFDTD_CHECK(hipMemcpyAsync(tmp, vx, sizeof(float) * st->Nvx, hipMemcpyHostToDevice, st->stream));
        if (ax == 1 && st->Nvy > 0)
            AFDTD_CHECK(upload_rows(tmp, vy, (size_t)(Ny - 1) * Nz, (size_t)Ny * Nz, Nx, st->stream));
        if (ax == 2 && st->Nvz > 0)
            AFDTD_CHECK(upload_rows(tmp, vz, (size_t)(Nz - 1), (size_t)Nz, (size_t)Nx * Ny, st->stream));
        rc = put(st->V[ax], st->sp.sV);
        AFDTD_CHECK(hipStreamSynchronize(st->stream));
    }
    AFDTD_CHECK(hipStreamSynchronize(st->stream));
    dfree(tmp);
    return rc;
}

}  // extern "C"
