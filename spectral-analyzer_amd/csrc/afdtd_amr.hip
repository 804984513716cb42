// afdtd_amr.hip — the topology-driven (AMR) acoustic stepper of libafdtd.so:
// the reference's amr_* C ABI (csrc/include/acoustic_amr.h) for gradient
// order 2, on the GPU.
//
// Device layout (HBM): per-face and per-cell records are one 128-bit load,
// a CSR entry is one 32-bit index whose sign selects the weight's sign:
//   face[f]  int4   {neg cell, pos cell, bits(1/d), bits(V_damp)}
//   v[f]     float  face velocity (neg -> pos)
//   p[c]     float  cell pressure
//   cell[c]  float4 {P_damp, 1/(V*open), P_src, bits(kind)}
//                   kind bit 0: wall/plate cell (zeroed after the update)
//                   kind bit 1: acoustic cell (injection candidate)
//   start[c] int    CSR row start (n_cells + 1)
//   csr[k]   int    face index f (the cell is the face's neg cell) or ~f (its
//                   pos cell), in face order per cell
//   w[f]     float  area*open; the CSR weight is +w[f] or -w[f]
// A step is two kernels on the handle's stream: k_amr_faces (one thread per
// face, fully coalesced record + gathered pressure pair) and k_amr_cells
// (one thread per cell: CSR walk, fused divergence + pressure update, wall
// zeroing and the stability test).  FP32 arithmetic follows the reference's
// expression order (acoustic_amr.cpp:1147-1172) and the library is compiled
// without FMA contraction, so results equal the reference's float path.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/afdtd_amr.h"

static thread_local std::string g_amr_stage = "ok";
static thread_local std::string g_amr_msg;

struct AcousticAMRState {
    int nc = 0, nf = 0, ne = 0, order = 2, step_count = 0;
    double c = 0.0, rho = 0.0, dt = 0.0;
    hipStream_t stream = nullptr;
    int4* d_face = nullptr;
    float* d_v = nullptr;
    float* d_p = nullptr;
    float4* d_cell = nullptr;
    int* d_start = nullptr;
    int* d_csr = nullptr;
    float* d_w = nullptr;
    double* d_cent = nullptr;
    int* d_status = nullptr;                 // [0] failing step, [1] nearest cell
    unsigned long long* d_best = nullptr;    // nearest squared distance (bits)
    // host copies (set-up inputs, coefficient rebuilds, getters)
    std::vector<double> centers, open_vol;
    std::vector<uint8_t> type;
    std::vector<int> neg, pos;
    std::vector<float> inv_d, inv_den, pdamp, psrc, vdamp;
    std::vector<int> kind;
};

namespace {

void set_stage(const char* s, const char* m) { g_amr_stage = s; g_amr_msg = m; }

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8);
// give each XCD a contiguous range of tiles instead, so the neighbour
// pressures and velocities one tile gathers were most likely brought into
// that XCD's L2 by the tiles just before it.
__device__ inline int xcd_tile(int nblk) {
    const int b = blockIdx.x, x = b & 7, i = b >> 3;
    const int q = nblk >> 3, r = nblk & 7;           // XCD x owns q + (x < r) tiles
    return x * q + min(x, r) + i;
}

constexpr int kFaceTile = 512;                       // faces per 256-thread block

__global__ __launch_bounds__(256) void k_amr_faces(const int4* __restrict__ face,
                                                   float* __restrict__ v,
                                                   const float* __restrict__ p, int nf,
                                                   float dt_over_rho,
                                                   const int* __restrict__ status, int s) {
    if (status[0] < s) return;                       // an earlier step of this call failed
    const int f0 = xcd_tile(gridDim.x) * kFaceTile + threadIdx.x;
    const int f1 = f0 + 256;
    // both faces' loads are issued before either update
    int4 r0 = make_int4(0, 0, 0, 0), r1 = r0;
    float v0 = 0.0f, v1 = 0.0f;
    if (f0 < nf) { r0 = face[f0]; v0 = v[f0]; }
    if (f1 < nf) { r1 = face[f1]; v1 = v[f1]; }
    const float a0 = p[r0.y], b0 = p[r0.x], a1 = p[r1.y], b1 = p[r1.x];
    if (f0 < nf) {
        const float g = (a0 - b0) * __int_as_float(r0.z);            // :1149-1150
        v[f0] = (v0 - dt_over_rho * g) * __int_as_float(r0.w);        // :1151
    }
    if (f1 < nf) {
        const float g = (a1 - b1) * __int_as_float(r1.z);
        v[f1] = (v1 - dt_over_rho * g) * __int_as_float(r1.w);
    }
}

// One 256-thread block updates 256 consecutive cells.  The block's CSR rows
// are one contiguous entry range; when it fits kCsrStage entries it is first
// copied into LDS with coalesced loads, so each lane's walk over its own row
// (about six entries on a box grid, 24 bytes apart between lanes) reads LDS
// instead of scattering loads over HBM lines.  Negating the face weight is
// exact, so +-w[f] equals the reference's pre-signed CSR weight.
constexpr int kCsrStage = 2048;

__global__ __launch_bounds__(256) void k_amr_cells(const float4* __restrict__ cell,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ csr,
                                                   const float* __restrict__ w,
                                                   const float* __restrict__ v,
                                                   float* __restrict__ p, int nc, float bulk,
                                                   int check_max, int* status, int s) {
    __shared__ int s_csr[kCsrStage];
    if (status[0] < s) return;
    const int c0 = xcd_tile(gridDim.x) * 256;
    const int c = c0 + threadIdx.x;
    const int k_lo = start[c0];
    const int k_hi = start[min(c0 + 256, nc)];
    const bool staged = (k_hi - k_lo) <= kCsrStage;  // uniform over the block
    if (staged) {
        for (int k = k_lo + threadIdx.x; k < k_hi; k += 256) s_csr[k - k_lo] = csr[k];
        __syncthreads();
    }
    bool bad = false;
    if (c < nc) {
        const float4 m = cell[c];
        const int k0 = start[c], k1 = start[c + 1];
        float d = 0.0f;
        if (staged) {
            for (int k = k0; k < k1; ++k) {                          // :1166-1168
                const int e = s_csr[k - k_lo];
                const int f = e < 0 ? ~e : e;
                d += (e < 0 ? -w[f] : w[f]) * v[f];
            }
        } else {
            for (int k = k0; k < k1; ++k) {
                const int e = csr[k];
                const int f = e < 0 ? ~e : e;
                d += (e < 0 ? -w[f] : w[f]) * v[f];
            }
        }
        float pn = p[c] * m.x - bulk * d * m.y * m.z;                 // :1169-1170
        if (__float_as_int(m.w) & 1) pn = 0.0f;                       // :1174-1176
        p[c] = pn;
        bad = !isfinite(pn) || (check_max && fabsf(pn) > 2e3f);       // :1181-1192
    }
    if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicMin(&status[0], s);
}

__global__ void k_amr_status_init(int* status, unsigned long long* best) {
    status[0] = INT_MAX;
    status[1] = INT_MAX;
    best[0] = ~0ull;
}

__global__ __launch_bounds__(256) void k_amr_near_d2(const double* __restrict__ cent,
                                                     const float4* __restrict__ cell, int nc,
                                                     double x, double y, double z,
                                                     unsigned long long* best) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    unsigned long long key = ~0ull;
    if (c < nc && (__float_as_int(cell[c].w) & 2)) {
        const double dx = cent[3 * (size_t)c] - x;
        const double dy = cent[3 * (size_t)c + 1] - y;
        const double dz = cent[3 * (size_t)c + 2] - z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        // the reference keeps a candidate only if d2 < 1e300 (its initial best)
        if (d2 < 1e300) key = (unsigned long long)__double_as_longlong(d2);
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long k2 = __shfl_xor(key, o);
        key = k2 < key ? k2 : key;
    }
    if ((threadIdx.x & 63) == 0 && key != ~0ull) atomicMin(best, key);
}

__global__ __launch_bounds__(256) void k_amr_near_idx(const double* __restrict__ cent,
                                                      const float4* __restrict__ cell, int nc,
                                                      double x, double y, double z,
                                                      const unsigned long long* best,
                                                      int* status) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= nc || !(__float_as_int(cell[c].w) & 2)) return;
    const double dx = cent[3 * (size_t)c] - x;
    const double dy = cent[3 * (size_t)c + 1] - y;
    const double dz = cent[3 * (size_t)c + 2] - z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < 1e300 && (unsigned long long)__double_as_longlong(d2) == best[0])
        atomicMin(&status[1], c);                    // first in index order on ties
}

__global__ void k_amr_near_add(float* p, const int* status, float value) {
    const int i = status[1];
    if (i != INT_MAX) p[i] += value;                 // :1242
}

__global__ __launch_bounds__(256) void k_fill_i32(int* a, size_t n, int val) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = val;
}

__global__ __launch_bounds__(256) void k_amr_scatter_idx(const double* __restrict__ cent, int nc,
                                                         int Nx, int Ny, int Nz, double x0,
                                                         double y0, double z0, double rx,
                                                         double ry, double rz, int* vox) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= nc) return;
    const int ix = (int)lround((cent[3 * (size_t)c] - x0) * rx);      // :1286-1288
    const int iy = (int)lround((cent[3 * (size_t)c + 1] - y0) * ry);
    const int iz = (int)lround((cent[3 * (size_t)c + 2] - z0) * rz);
    if (ix < 0 || ix >= Nx || iy < 0 || iy >= Ny || iz < 0 || iz >= Nz) return;
    atomicMax(&vox[(size_t)ix * Ny * Nz + (size_t)iy * Nz + iz], c);  // last write wins
}

__global__ __launch_bounds__(256) void k_amr_scatter_val(const int* __restrict__ vox,
                                                         const float* __restrict__ p,
                                                         float* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int c = vox[i];
    out[i] = c >= 0 ? p[c] : 0.0f;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

int validate(int nc, const double* cc, const double* vol, const double* ov, const uint8_t* ct,
             int nf, const int32_t* fn, const int32_t* fp, const double* fa, const double* fo,
             const double* fd, double c, double rho, double min_dx) {
    // acoustic_amr.cpp:275-311
    if (nc <= 0 || nf <= 0) return SK_ERR_DIM_MISMATCH;
    if (!cc || !vol || !ov || !ct || !fn || !fp || !fa || !fo || !fd) return SK_ERR_NULL_STATE;
    if (c <= 0.0 || rho <= 0.0 || min_dx <= 0.0) return SK_ERR_DIM_MISMATCH;
    for (int i = 0; i < nc; ++i) {
        if (!(vol[i] > 0.0) || ov[i] < 0.0 || ov[i] > 1.0 || ct[i] > 3) return SK_ERR_DIM_MISMATCH;
    }
    for (int f = 0; f < nf; ++f) {
        if (fn[f] < 0 || fn[f] >= nc || fp[f] < 0 || fp[f] >= nc || fn[f] == fp[f])
            return SK_ERR_DIM_MISMATCH;
        if (!(fa[f] > 0.0) || fo[f] < 0.0 || fo[f] > 1.0 || !(fd[f] > 0.0))
            return SK_ERR_DIM_MISMATCH;
    }
    return SK_OK;
}

void destroy_state(AcousticAMRState* st) {
    if (!st) return;
    if (st->stream) (void)hipStreamSynchronize(st->stream);
    void* bufs[] = {st->d_face, st->d_v, st->d_p, st->d_cell, st->d_start,
                    st->d_csr,  st->d_w,    st->d_cent, st->d_status, st->d_best};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (st->stream) (void)hipStreamDestroy(st->stream);
    delete st;
}

// Pack the per-cell and per-face coefficient records and upload them.
hipError_t upload_coeffs(AcousticAMRState* st) {
    std::vector<float4> cell(st->nc);
    for (int i = 0; i < st->nc; ++i) {
        float kf;
        const int k = st->kind[i];
        memcpy(&kf, &k, 4);
        cell[i] = make_float4(st->pdamp[i], st->inv_den[i], st->psrc[i], kf);
    }
    std::vector<int4> face(st->nf);
    for (int f = 0; f < st->nf; ++f) {
        int a, b;
        memcpy(&a, &st->inv_d[f], 4);
        memcpy(&b, &st->vdamp[f], 4);
        face[f] = make_int4(st->neg[f], st->pos[f], a, b);
    }
    hipError_t e = hipMemcpyAsync(st->d_cell, cell.data(), sizeof(float4) * st->nc,
                                  hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(st->d_face, face.data(), sizeof(int4) * st->nf, hipMemcpyHostToDevice,
                           st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    return e;
}

void report_diverged(AcousticAMRState* st) {
    // acoustic_amr.cpp:1074-1093: one stderr line with the worst cell
    std::vector<float> p(st->nc);
    if (hipMemcpy(p.data(), st->d_p, sizeof(float) * st->nc, hipMemcpyDeviceToHost) != hipSuccess)
        return;
    int worst = 0, n_bad = 0;
    float wp = 0.0f;
    for (int i = 0; i < st->nc; ++i) {
        if (!isfinite(p[i])) ++n_bad;
        if (fabsf(p[i]) > fabsf(wp)) { wp = p[i]; worst = i; }
    }
    fprintf(stderr,
            "[amr_step] UNSTABLE step=%d p_max=%.3g Pa worst_cell=%d pos=(%.4f, %.4f, %.4f) "
            "n_nonfinite=%d\n",
            st->step_count, wp, worst, (float)st->centers[3 * (size_t)worst],
            (float)st->centers[3 * (size_t)worst + 1], (float)st->centers[3 * (size_t)worst + 2],
            n_bad);
}

int step_impl(AcousticAMRState* st, int n, float* ms_total, float* ms_faces, float* ms_cells) {
    if (!st) return SK_ERR_NULL_STATE;
    if (n < 0) return SK_ERR_DIM_MISMATCH;
    if (n == 0) {
        if (ms_total) *ms_total = *ms_faces = *ms_cells = 0.0f;
        return SK_OK;
    }
    const float dt = (float)st->dt;                                  // :1111-1114
    const float rho = (float)st->rho;
    const float bulk = (float)(st->rho * st->c * st->c * st->dt);
    const float dt_over_rho = dt / rho;
    std::vector<hipEvent_t> ev;
    if (ms_total) {
        ev.resize(2 * (size_t)n + 1);
        for (auto& e : ev)
            if (hipEventCreate(&e) != hipSuccess) return AMR_ERR_DEVICE;
    }
    k_amr_status_init<<<1, 1, 0, st->stream>>>(st->d_status, st->d_best);
    for (int s = 0; s < n; ++s) {
        if (ms_total) (void)hipEventRecord(ev[2 * (size_t)s], st->stream);
        k_amr_faces<<<(st->nf + kFaceTile - 1) / kFaceTile, 256, 0, st->stream>>>(
            st->d_face, st->d_v, st->d_p, st->nf, dt_over_rho, st->d_status, s);
        if (ms_total) (void)hipEventRecord(ev[2 * (size_t)s + 1], st->stream);
        const int check_max = ((st->step_count + s) & 0xFF) == 0;
        k_amr_cells<<<blocks(st->nc), 256, 0, st->stream>>>(st->d_cell, st->d_start, st->d_csr,
                                                          st->d_w, st->d_v, st->d_p, st->nc, bulk,
                                                          check_max, st->d_status, s);
    }
    if (ms_total) (void)hipEventRecord(ev[2 * (size_t)n], st->stream);
    int fail = INT_MAX;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(&fail, st->d_status, sizeof(int), hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    if (ms_total && e == hipSuccess) {
        float tf = 0.0f, tc = 0.0f, x = 0.0f;
        for (int s = 0; s < n; ++s) {
            (void)hipEventElapsedTime(&x, ev[2 * (size_t)s], ev[2 * (size_t)s + 1]);
            tf += x;
            (void)hipEventElapsedTime(&x, ev[2 * (size_t)s + 1], ev[2 * (size_t)s + 2]);
            tc += x;
        }
        (void)hipEventElapsedTime(&x, ev[0], ev[2 * (size_t)n]);
        *ms_total = x;
        *ms_faces = tf;
        *ms_cells = tc;
    }
    for (auto& v : ev) (void)hipEventDestroy(v);
    if (e != hipSuccess) return AMR_ERR_DEVICE;
    if (fail < n) {                                  // steps before it completed
        st->step_count += fail;
        report_diverged(st);
        return SK_ERR_DIVERGED;
    }
    st->step_count += n;
    return SK_OK;
}

}  // namespace

extern "C" {

SK_API AcousticAMRState* amr_create(int nc, const double* cc, const double* vol, const double* ov,
                                    const uint8_t* ct, int nf, const int32_t* fn,
                                    const int32_t* fp, const double* fa, const double* fo,
                                    const double* fd, double c, double rho, double min_dx,
                                    int order) {
    set_stage("ok", "");
    if (validate(nc, cc, vol, ov, ct, nf, fn, fp, fa, fo, fd, c, rho, min_dx) != SK_OK) {
        set_stage("validate_inputs", "input validation failed");
        return nullptr;
    }
    if (order != 2 && order != 8) {
        set_stage("validate_inputs", "gradient_order must be 2 or 8");
        return nullptr;
    }
    if (order == 8) {
        set_stage("validate_inputs",
                  "gradient_order 8 (Fornberg stencil) is not provided by this library");
        return nullptr;
    }
    auto* st = new AcousticAMRState();
    st->nc = nc;
    st->nf = nf;
    st->order = order;
    st->c = c;
    st->rho = rho;
    st->dt = 0.77 * min_dx / (c * sqrt(3.0));                        // :744-745
    st->centers.assign(cc, cc + 3 * (size_t)nc);
    st->open_vol.assign(ov, ov + nc);
    st->type.assign(ct, ct + nc);
    st->neg.assign(fn, fn + nf);
    st->pos.assign(fp, fp + nf);
    st->kind.assign(nc, 0);
    int n_acoustic = 0;
    for (int i = 0; i < nc; ++i) {
        if ((ct[i] == 0 || ct[i] == 3) && ov[i] > 0.0) { st->kind[i] |= 2; ++n_acoustic; }  // :783-785
        if (ct[i] == 1 || ct[i] == 2) st->kind[i] |= 1;                                   // :813-816
    }
    if (n_acoustic == 0) {
        set_stage("acoustic_cells", "no active acoustic cells after filtering");
        delete st;
        return nullptr;
    }
    st->inv_d.resize(nf);
    std::vector<float> ffc(nf);
    for (int f = 0; f < nf; ++f) {                                   // :801-804
        st->inv_d[f] = 1.0f / (float)fd[f];
        ffc[f] = (float)(fa[f] * fo[f]);
    }
    st->inv_den.assign(nc, 0.0f);
    for (int i = 0; i < nc; ++i)                                     // :808-811
        if (st->kind[i] & 2) st->inv_den[i] = (float)(1.0 / (vol[i] * std::max(ov[i], 1e-12)));
    std::vector<int> start(nc + 1, 0);                               // :818-831
    for (int f = 0; f < nf; ++f) { start[fn[f] + 1]++; start[fp[f] + 1]++; }
    for (int i = 1; i <= nc; ++i) start[i] += start[i - 1];
    st->ne = start[nc];
    std::vector<int> csr(st->ne);
    {
        std::vector<int> fill(nc, 0);
        for (int f = 0; f < nf; ++f) {                               // :837-848
            csr[start[fn[f]] + fill[fn[f]]++] = f;           // weight +ffc
            csr[start[fp[f]] + fill[fp[f]]++] = ~f;          // weight -ffc
        }
    }
    st->pdamp.assign(nc, 1.0f);                                      // :852-859
    st->psrc.assign(nc, 1.0f);
    st->vdamp.assign(nf, 1.0f);

    hipError_t e = hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking);
    auto al = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes ? bytes : 1);
    };
    al((void**)&st->d_face, sizeof(int4) * (size_t)nf);
    al((void**)&st->d_v, sizeof(float) * (size_t)nf);
    al((void**)&st->d_p, sizeof(float) * (size_t)nc);
    al((void**)&st->d_cell, sizeof(float4) * (size_t)nc);
    al((void**)&st->d_start, sizeof(int) * ((size_t)nc + 1));
    al((void**)&st->d_csr, sizeof(int) * (size_t)st->ne);
    al((void**)&st->d_w, sizeof(float) * (size_t)nf);
    al((void**)&st->d_cent, sizeof(double) * 3 * (size_t)nc);
    al((void**)&st->d_status, sizeof(int) * 2);
    al((void**)&st->d_best, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemsetAsync(st->d_v, 0, sizeof(float) * (size_t)nf, st->stream);
    if (e == hipSuccess) e = hipMemsetAsync(st->d_p, 0, sizeof(float) * (size_t)nc, st->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(st->d_start, start.data(), sizeof(int) * ((size_t)nc + 1),
                           hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(st->d_csr, csr.data(), sizeof(int) * (size_t)st->ne,
                           hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(st->d_w, ffc.data(), sizeof(float) * (size_t)nf,
                           hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(st->d_cent, cc, sizeof(double) * 3 * (size_t)nc,
                           hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess) e = upload_coeffs(st);
    if (e != hipSuccess) {
        set_stage("device_alloc", hipGetErrorString(e));
        destroy_state(st);
        return nullptr;
    }
    return st;
}

SK_API const char* amr_get_last_create_error_stage(void) { return g_amr_stage.c_str(); }
SK_API const char* amr_get_last_create_error_message(void) { return g_amr_msg.c_str(); }
SK_API int amr_get_create_progress_active(void) { return 0; }
SK_API int amr_get_create_progress_done_faces(void) { return 0; }
SK_API int amr_get_create_progress_total_faces(void) { return 0; }

SK_API void amr_destroy(AcousticAMRState* st) { destroy_state(st); }

SK_API int amr_step(AcousticAMRState* st, int n_steps) {
    return step_impl(st, n_steps, nullptr, nullptr, nullptr);
}

SK_API int amr_step_timed(AcousticAMRState* st, int n_steps, float* ms_total, float* ms_faces,
                          float* ms_cells) {
    if (!st || !ms_total || !ms_faces || !ms_cells) return SK_ERR_NULL_STATE;
    return step_impl(st, n_steps, ms_total, ms_faces, ms_cells);
}

SK_API int amr_reset(AcousticAMRState* st) {
    if (!st) return SK_ERR_NULL_STATE;                               // :924-936
    hipError_t e = hipMemsetAsync(st->d_p, 0, sizeof(float) * (size_t)st->nc, st->stream);
    if (e == hipSuccess) e = hipMemsetAsync(st->d_v, 0, sizeof(float) * (size_t)st->nf, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    st->step_count = 0;
    return e == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

SK_API int amr_set_state(AcousticAMRState* st, const float* p, const float* v) {
    if (!st) return SK_ERR_NULL_STATE;
    hipError_t e = hipSuccess;
    if (p)
        e = hipMemcpyAsync(st->d_p, p, sizeof(float) * (size_t)st->nc, hipMemcpyHostToDevice,
                           st->stream);
    if (v && e == hipSuccess)
        e = hipMemcpyAsync(st->d_v, v, sizeof(float) * (size_t)st->nf, hipMemcpyHostToDevice,
                           st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    return e == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

SK_API int amr_inject_pressure_nearest(AcousticAMRState* st, const double* xyz, float value) {
    if (!st || !xyz) return SK_ERR_NULL_STATE;                       // :1227-1244
    if (!isfinite(value)) return SK_ERR_DIVERGED;
    const double x = xyz[0], y = xyz[1], z = xyz[2];
    k_amr_status_init<<<1, 1, 0, st->stream>>>(st->d_status, st->d_best);
    k_amr_near_d2<<<blocks(st->nc), 256, 0, st->stream>>>(st->d_cent, st->d_cell, st->nc, x, y, z,
                                                        st->d_best);
    k_amr_near_idx<<<blocks(st->nc), 256, 0, st->stream>>>(st->d_cent, st->d_cell, st->nc, x, y,
                                                         z, st->d_best, st->d_status);
    k_amr_near_add<<<1, 1, 0, st->stream>>>(st->d_p, st->d_status, value);
    int idx = INT_MAX;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(&idx, st->d_status + 1, sizeof(int), hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    if (e != hipSuccess) return AMR_ERR_DEVICE;
    return idx == INT_MAX ? SK_ERR_DIM_MISMATCH : SK_OK;
}

SK_API int amr_get_pressure(const AcousticAMRState* st, float* out, int n) {
    if (!st || !out) return SK_ERR_NULL_STATE;                       // :1246-1252
    if (n != st->nc) return SK_ERR_DIM_MISMATCH;
    hipError_t e = hipMemcpyAsync(out, st->d_p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost,
                                  st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    return e == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

SK_API int amr_get_velocity(const AcousticAMRState* st, float* out, int n) {
    if (!st || !out) return SK_ERR_NULL_STATE;                       // :1296-1302
    if (n != st->nf) return SK_ERR_DIM_MISMATCH;
    hipError_t e = hipMemcpyAsync(out, st->d_v, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost,
                                  st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    return e == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

SK_API int amr_get_cell_centers(const AcousticAMRState* st, float* out, int n3) {
    if (!st || !out) return SK_ERR_NULL_STATE;                       // :1254-1264
    if (n3 < st->nc * 3) return SK_ERR_DIM_MISMATCH;
    for (size_t i = 0; i < 3 * (size_t)st->nc; ++i) out[i] = (float)st->centers[i];
    return SK_OK;
}

SK_API double amr_get_dt(const AcousticAMRState* st) { return st ? st->dt : 0.0; }
SK_API int amr_get_step_count(const AcousticAMRState* st) { return st ? st->step_count : 0; }
SK_API int amr_get_csr_entries(const AcousticAMRState* st) { return st ? st->ne : 0; }

SK_API int amr_scatter_pressure_uniform(const AcousticAMRState* st, int Nx, int Ny, int Nz,
                                        const double bmin[3], const double bmax[3], float* out) {
    if (!st || !out || !bmin || !bmax) return SK_ERR_NULL_STATE;     // :1266-1294
    if (Nx <= 0 || Ny <= 0 || Nz <= 0) return SK_ERR_DIM_MISMATCH;
    const size_t n = (size_t)Nx * Ny * Nz;
    const double rx = (Nx > 1) ? (Nx - 1) / (bmax[0] - bmin[0]) : 0.0;
    const double ry = (Ny > 1) ? (Ny - 1) / (bmax[1] - bmin[1]) : 0.0;
    const double rz = (Nz > 1) ? (Nz - 1) / (bmax[2] - bmin[2]) : 0.0;
    int* vox = nullptr;
    float* dout = nullptr;
    hipError_t e = hipMalloc((void**)&vox, sizeof(int) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&dout, sizeof(float) * n);
    if (e == hipSuccess) {
        k_fill_i32<<<blocks(n), 256, 0, st->stream>>>(vox, n, -1);
        k_amr_scatter_idx<<<blocks(st->nc), 256, 0, st->stream>>>(
            st->d_cent, st->nc, Nx, Ny, Nz, bmin[0], bmin[1], bmin[2], rx, ry, rz, vox);
        k_amr_scatter_val<<<blocks(n), 256, 0, st->stream>>>(vox, st->d_p, dout, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, sizeof(float) * n, hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    if (vox) (void)hipFree(vox);
    if (dout) (void)hipFree(dout);
    return e == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

SK_API int amr_set_border_condition(AcousticAMRState* st, int mode, float sigma_order,
                                    float R_reflection, float border_Z_match, int n_pml,
                                    const double* bmin, const double* bmax) {
    (void)border_Z_match;                            // read by no mode (:1318-1409)
    if (!st || !bmin || !bmax) return SK_ERR_NULL_STATE;
    if (n_pml <= 0) return SK_OK;
    const float order = (sigma_order > 0.0f) ? sigma_order : 3.0f;
    const float dt = (float)st->dt;
    float smax = 3.45f / (dt * (float)n_pml);
    if (mode == AMR_BORDER_ROOM_PANEL) smax *= (1.0f - std::max(0.0f, std::min(1.0f, R_reflection)));
    // min_dx recovered with the order-2 Courant number, as the reference does
    const double depth = (double)n_pml * (st->dt * st->c * sqrt(3.0) / 0.77);
    std::vector<float> alpha(st->nc, 0.0f);
    for (int i = 0; i < st->nc; ++i) {
        if (st->type[i] != 3) continue;
        const double* q = &st->centers[3 * (size_t)i];
        const double d = std::max(std::min({q[0] - bmin[0], bmax[0] - q[0], q[1] - bmin[1],
                                            bmax[1] - q[1], q[2] - bmin[2], bmax[2] - q[2]}),
                                  0.0);
        const float t = (float)std::max(0.0, std::min(1.0, 1.0 - d / depth));
        alpha[i] = smax * powf(t, order);
    }
    for (int i = 0; i < st->nc; ++i) {
        const float s = alpha[i] * dt;
        st->pdamp[i] = expf(-s);
        st->psrc[i] = (s < 1e-6f) ? 1.0f - 0.5f * s : (1.0f - expf(-s)) / s;
    }
    for (int f = 0; f < st->nf; ++f) {
        const int a = st->neg[f], b = st->pos[f];
        const float vn = (float)st->open_vol[a], vp = (float)st->open_vol[b];
        const float den = vn + vp;
        const float sf = (den > 0.f) ? 2.0f * (vp * alpha[a] + vn * alpha[b]) / den
                                     : 0.5f * (alpha[a] + alpha[b]);
        st->vdamp[f] = expf(-sf * dt);
    }
    return upload_coeffs(st) == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

}  // extern "C"
