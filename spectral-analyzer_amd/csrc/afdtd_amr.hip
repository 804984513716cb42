// afdtd_amr.hip — the topology-driven (AMR) acoustic stepper of libafdtd.so:
// the reference's amr_* C ABI (csrc/include/acoustic_amr.h) on the GPU,
// gradient orders 2 and 8, border conditions, Kirchhoff plate, mic samplers.
//
// Device layout (HBM), direct records (every record streamed in full):
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
//   sc/sw           order 8: 8 stencil cells and weights per face
// Compact records (used while the tables stay small): per face {neg, class},
// per cell a class id, into tables of the distinct records.
// A step is two kernels on the handle's stream, the face update
// (k_amr_faces / _faces_c / _faces8 / _faces8_c) and the cell update
// (k_amr_cells: CSR walk, fused divergence + pressure update, wall zeroing,
// stability test), plus the plate's two kernels when a plate is attached.
// FP32 arithmetic follows the reference's expression order
// (acoustic_amr.cpp:1127-1172) and the library is compiled without FMA
// contraction, so results equal the reference's float path.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <unordered_map>
#include <string>
#include <vector>

#include "../../include/afdtd_amr.h"
#include "afdtd_internal.h"

static thread_local std::string g_amr_stage = "ok";
// stencil-build progress of the last order-8 create (global, as the reference's)
static std::atomic<int> g_done_faces{0}, g_total_faces{0};
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
    int4* d_sc = nullptr;                    // order 8: 8 cells per face (two int4)
    float4* d_sw = nullptr;                  // order 8: 8 weights per face (two float4)
    std::vector<int> sc8;                    // order 8 stencils, kept for record rebuilds
    std::vector<float> sw8;
    // Compact records: each face stores {neg cell, class} and each cell a
    // class id into small tables of the distinct records (a box grid has a
    // handful), cutting the streamed bytes; used when the tables stay small.
    int layout_req = 0;                      // 0 automatic, 1 direct, 2 compact
    bool compact = false;
    int n_fclass = 0, n_cclass = 0;
    int2* d_fc = nullptr;                    // per face {neg, class}
    int4* d_fcls = nullptr;                  // order 2: {pos-neg, 1/d, Vdamp, 0}; order 8: 5 per class
    int* d_cc = nullptr;                     // per cell class
    float4* d_ccls = nullptr;                // {P_damp, 1/(V open), P_src, kind}
    int* d_status = nullptr;                 // [0] failing step, [1] nearest cell
    unsigned long long* d_best = nullptr;    // nearest squared distance (bits)
    // host copies (set-up inputs, coefficient rebuilds, getters)
    std::vector<double> centers, open_vol;
    std::vector<uint8_t> type;
    std::vector<int> neg, pos;
    std::vector<float> inv_d, inv_den, pdamp, psrc, vdamp;
    std::vector<int> kind;
    std::vector<double> volumes, face_area, face_open;
    // Kirchhoff plate: three rotating displacement buffers (w_prev, w, w_new
    // at rotation prot), external force, per-node face lists (CSR over the
    // active nodes) and the faces the plate drives (one writer per face)
    bool plate_on = false;
    int pNx = 0, pNy = 0, n_pnodes = 0, n_active = 0, prot = 0, n_bc = 0;
    float p_dx = 0, p_rh = 0, p_D = 0, p_aM = 0, p_bK = 0;
    float* d_pw[3] = {nullptr, nullptr, nullptr};
    float* d_ext = nullptr;
    int* d_pint = nullptr;     // active[n_active], ab_start/bl_start[n_active+1],
                               // ab_side[la], bl_side[lb], cabove/cbelow[n_active],
                               // bc_face/bc_node/bc_sign[n_bc]
    float* d_pflt = nullptr;   // ab_wgt[la], bl_wgt[lb]
    int o_active = 0, o_ab = 0, o_bl = 0, o_abs = 0, o_bls = 0, o_ca = 0, o_cb = 0;
    int o_bcf = 0, o_bcn = 0, o_bcs = 0, o_abw = 0, o_blw = 0;
    std::vector<int> bridge_idx, neck_idx;
    std::vector<float> bridge_wgt, neck_wgt;
    // mic samplers: growing per-mic lists as the reference keeps them, and
    // their CSR copy on the device
    std::vector<std::vector<int>> mp_cell, mv_face;
    std::vector<std::vector<float>> mp_w, mv_wx, mv_wy, mv_wz;
    int n_mics = 0;
    int* d_mint = nullptr;     // mp_start[n+1], mv_start[n+1], cells, faces
    float* d_mflt = nullptr;   // p weights, then wx, wy, wz
    float* d_mout = nullptr;   // 4*n_mics outputs
    int o_mvs = 0, o_mpc = 0, o_mvf = 0, o_mwx = 0, o_mwy = 0, o_mwz = 0;
    // small upload buffers (drives, forces, sampler stencils)
    void* d_tmp = nullptr;
    size_t tmp_bytes = 0;
    // Captured launch sequences of a block of recorded audio samples, replayed
    // while their key (block shape, device pointers and coefficients via
    // `gen`) is unchanged.  The plate's three displacement buffers rotate by
    // the block's step count, so a plated grid keeps one graph per rotation
    // (slot prot % 3): blocks whose step count is not a multiple of 3 cycle
    // through the slots instead of capturing anew every call.
    uint64_t gen = 0;                        // bumped whenever captured arguments change
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    std::array<uint64_t, 6> gkey[3]{};
};

namespace {

void set_stage(const char* s, const char* m) { g_amr_stage = s; g_amr_msg = m; }

// Drops the captured block graph.  Called before any buffer a captured
// launch references is freed or replaced, so that no failure path between
// the free and the bump of `gen` can leave a graph pointing at freed memory.
void drop_graph(AcousticAMRState* st) {
    for (auto& g : st->gexec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8);
// give each XCD a contiguous range of tiles instead, so the neighbour
// pressures and velocities one tile gathers were most likely brought into
// that XCD's L2 by the tiles just before it.
__device__ inline int xcd_tile(int nblk) {
    const int b = blockIdx.x, x = b & 7, i = b >> 3;
    const int q = nblk >> 3, r = nblk & 7;           // XCD x owns q + (x < r) tiles
    return x * q + min(x, r) + i;
}

constexpr int kFacesPerLane = 2;                     // direct order-2 face kernel
constexpr int kFaceTileD = 256 * kFacesPerLane;

__global__ __launch_bounds__(256) void k_amr_faces(const int4* __restrict__ face,
                                                   float* __restrict__ v,
                                                   const float* __restrict__ p, int nf,
                                                   float dt_over_rho,
                                                   const int* __restrict__ status, int s) {
    if (status[0] < s) return;                       // an earlier step of this call failed
    constexpr int U = kFacesPerLane;
    const int fb = xcd_tile(gridDim.x) * kFaceTileD + threadIdx.x;
    // every face's loads are issued before any update
    int4 r[U];
    float vo[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int f = fb + 256 * u;
        r[u] = make_int4(0, 0, 0, 0);
        vo[u] = 0.0f;
        if (f < nf) { r[u] = face[f]; vo[u] = v[f]; }
    }
    float a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) { a[u] = p[r[u].y]; b[u] = p[r[u].x]; }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int f = fb + 256 * u;
        if (f < nf) {
            const float g = (a[u] - b[u]) * __int_as_float(r[u].z);    // :1149-1150
            v[f] = (vo[u] - dt_over_rho * g) * __int_as_float(r[u].w); // :1151
        }
    }
}

// Gradient order 8 (acoustic_amr.cpp:1127-1143): the face's 8-cell stencil
// (two 16-byte cell records, two 16-byte weight records per face), summed
// in slot order from 0 including the zero-weight padding slots (cell 0).
__global__ __launch_bounds__(256) void k_amr_faces8(const int4* __restrict__ sc,
                                                    const float4* __restrict__ sw,
                                                    const int4* __restrict__ face,
                                                    float* __restrict__ v,
                                                    const float* __restrict__ p, int nf,
                                                    float dt_over_rho,
                                                    const int* __restrict__ status, int s) {
    if (status[0] < s) return;
    const int f = xcd_tile(gridDim.x) * 256 + threadIdx.x;
    if (f >= nf) return;
    const int4 a = sc[2 * (size_t)f], b = sc[2 * (size_t)f + 1];
    const float4 wa = sw[2 * (size_t)f], wb = sw[2 * (size_t)f + 1];
    const float vd = __int_as_float(face[f].w);
    const float pa[8] = {p[a.x], p[a.y], p[a.z], p[a.w], p[b.x], p[b.y], p[b.z], p[b.w]};
    float g = 0.0f;
    g += wa.x * pa[0];
    g += wa.y * pa[1];
    g += wa.z * pa[2];
    g += wa.w * pa[3];
    g += wb.x * pa[4];
    g += wb.y * pa[5];
    g += wb.z * pa[6];
    g += wb.w * pa[7];
    v[f] = (v[f] - dt_over_rho * g) * vd;
}

// Compact-record variants of the face kernels: a face stores {neg, class};
// the class table (a few hundred bytes on a box grid, cache resident)
// holds the rest of the record, bit for bit the direct values.
constexpr int kFacesPerLaneC = 4;                    // compact order-2 face kernel
constexpr int kFaceTileC = 256 * kFacesPerLaneC;

__global__ __launch_bounds__(256) void k_amr_faces_c(const int2* __restrict__ fc,
                                                     const int4* __restrict__ cls,
                                                     float* __restrict__ v,
                                                     const float* __restrict__ p, int nf,
                                                     float dt_over_rho,
                                                     const int* __restrict__ status, int s) {
    if (status[0] < s) return;
    constexpr int U = kFacesPerLaneC;
    const int fb = xcd_tile(gridDim.x) * kFaceTileC + threadIdx.x;
    // every face's loads are issued before any update
    int2 r[U];
    float vo[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int f = fb + 256 * u;
        r[u] = make_int2(0, 0);
        vo[u] = 0.0f;
        if (f < nf) { r[u] = fc[f]; vo[u] = v[f]; }
    }
    int4 c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = cls[r[u].y];
    float a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        // lanes past the last face read cell 0 only (class 0's offset may be negative)
        const int q = fb + 256 * u < nf ? r[u].x + c[u].x : 0;
        a[u] = p[q];
        b[u] = p[r[u].x];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int f = fb + 256 * u;
        if (f < nf)
            v[f] = (vo[u] - dt_over_rho * ((a[u] - b[u]) * __int_as_float(c[u].y))) *
                   __int_as_float(c[u].z);
    }
}

__device__ inline int stencil_cell(int neg, int d) { return d == (int)0x80000000 ? 0 : neg + d; }

// Several faces per lane: every face's 8 gathers are issued before any sum.
__device__ inline float faces8_grad(const int4* __restrict__ c, int n,
                                    const float* __restrict__ p, float* vd) {
    const int4 a = c[0], b = c[1], wa = c[2], wb = c[3];
    *vd = __int_as_float(c[4].x);
    const float pa[8] = {p[stencil_cell(n, a.x)], p[stencil_cell(n, a.y)], p[stencil_cell(n, a.z)],
                         p[stencil_cell(n, a.w)], p[stencil_cell(n, b.x)], p[stencil_cell(n, b.y)],
                         p[stencil_cell(n, b.z)], p[stencil_cell(n, b.w)]};
    float g = 0.0f;
    g += __int_as_float(wa.x) * pa[0];
    g += __int_as_float(wa.y) * pa[1];
    g += __int_as_float(wa.z) * pa[2];
    g += __int_as_float(wa.w) * pa[3];
    g += __int_as_float(wb.x) * pa[4];
    g += __int_as_float(wb.y) * pa[5];
    g += __int_as_float(wb.z) * pa[6];
    g += __int_as_float(wb.w) * pa[7];
    return g;
}

constexpr int kFacesPerLane8C = 2;
constexpr int kFaceTile8C = 256 * kFacesPerLane8C;

__global__ __launch_bounds__(256) void k_amr_faces8_c(const int2* __restrict__ fc,
                                                      const int4* __restrict__ cls,
                                                      float* __restrict__ v,
                                                      const float* __restrict__ p, int nf,
                                                      float dt_over_rho,
                                                      const int* __restrict__ status, int s) {
    if (status[0] < s) return;
    constexpr int U = kFacesPerLane8C;
    const int fb = xcd_tile(gridDim.x) * kFaceTile8C + threadIdx.x;
    // lanes past the last face evaluate face 0's stencil (in bounds) and store nothing
    int2 r[U];
    float vo[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int f = fb + 256 * u;
        r[u] = fc[f < nf ? f : 0];
        vo[u] = f < nf ? v[f] : 0.0f;
    }
    float g[U], vd[U];
#pragma unroll
    for (int u = 0; u < U; ++u) g[u] = faces8_grad(cls + 5 * (size_t)r[u].y, r[u].x, p, &vd[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int f = fb + 256 * u;
        if (f < nf) v[f] = (vo[u] - dt_over_rho * g[u]) * vd[u];
    }
}

// One 256-thread block updates 512 consecutive cells.  The block's CSR rows
// are one contiguous entry range; when it fits kCsrStage entries it is first
// copied into LDS with coalesced loads, so each lane's walk over its own row
// (about six entries on a box grid, 24 bytes apart between lanes) reads LDS
// instead of scattering loads over HBM lines.  Negating the face weight is
// exact, so +-w[f] equals the reference's pre-signed CSR weight.
constexpr int kCellTile = 512;                       // cells per 256-thread block
constexpr int kRowMax = 8;                           // CSR rows walked unrolled up to this
constexpr int kCsrStage = 4096;

// Accumulate one CSR row entry (the reference's signed weight times v).
__device__ inline float csr_term(int e, const float* __restrict__ w,
                                 const float* __restrict__ v) {
    const int f = e < 0 ? ~e : e;
    return (e < 0 ? -w[f] : w[f]) * v[f];
}

// Each lane updates two cells, c and c + 256, walking both rows in step so
// twice as many gathers are in flight per lane (the kernel is bound by the
// latency of its gathers at full occupancy); each row is still summed in
// CSR order.
template <bool COMPACT>
__global__ __launch_bounds__(256) void k_amr_cells(const float4* __restrict__ cell,
                                                   const int* __restrict__ cell_class,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ csr,
                                                   const float* __restrict__ w,
                                                   const float* __restrict__ v,
                                                   float* __restrict__ p, int nc, float bulk,
                                                   int* status, int s) {
    __shared__ int s_csr[kCsrStage];
    if (status[0] < s) return;
    // |p| is tested every 256th step of the handle (status[2]: its step count)
    const int check_max = ((status[2] + s) & 0xFF) == 0;
    const int c0 = xcd_tile(gridDim.x) * kCellTile;
    const int k_lo = start[c0];
    const int k_hi = start[min(c0 + kCellTile, nc)];
    const bool staged = (k_hi - k_lo) <= kCsrStage;  // uniform over the block
    if (staged) {
        for (int k = k_lo + threadIdx.x; k < k_hi; k += 256) s_csr[k - k_lo] = csr[k];
        __syncthreads();
    }
    const int ca = c0 + threadIdx.x, cb = ca + 256;
    const bool va = ca < nc, vb = cb < nc;
    const int a0 = va ? start[ca] : 0, a1 = va ? start[ca + 1] : 0;
    const int b0 = vb ? start[cb] : 0, b1 = vb ? start[cb + 1] : 0;
    const int na = a1 - a0, nb = b1 - b0;
    // the cells' own records and pressures load alongside the row gathers
    float4 ma = make_float4(0.0f, 0.0f, 0.0f, 0.0f), mb = ma;
    float pa = 0.0f, pb = 0.0f;
    if (va) { ma = COMPACT ? cell[cell_class[ca]] : cell[ca]; pa = p[ca]; }
    if (vb) { mb = COMPACT ? cell[cell_class[cb]] : cell[cb]; pb = p[cb]; }
    float da = 0.0f, db = 0.0f;
    if (staged && na <= kRowMax && nb <= kRowMax) {
        // rows of up to kRowMax faces (six on a box grid): every gather of
        // both rows is issued before the first add, then each row is summed
        // in CSR order (:1166-1168); on small grids, where a launch is a
        // chain of dependent loads, this shortens the chain
        float ta[kRowMax], tb[kRowMax];
#pragma unroll
        for (int j = 0; j < kRowMax; ++j) {
            ta[j] = j < na ? csr_term(s_csr[a0 - k_lo + j], w, v) : 0.0f;
            tb[j] = j < nb ? csr_term(s_csr[b0 - k_lo + j], w, v) : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < kRowMax; ++j) {
            if (j < na) da += ta[j];
            if (j < nb) db += tb[j];
        }
    } else if (staged) {
        for (int j = 0; j < max(na, nb); ++j) {                      // :1166-1168
            if (j < na) da += csr_term(s_csr[a0 - k_lo + j], w, v);
            if (j < nb) db += csr_term(s_csr[b0 - k_lo + j], w, v);
        }
    } else {
        for (int j = 0; j < max(na, nb); ++j) {
            if (j < na) da += csr_term(csr[a0 + j], w, v);
            if (j < nb) db += csr_term(csr[b0 + j], w, v);
        }
    }
    bool bad = false;
    if (va) {
        float pn = pa * ma.x - bulk * da * ma.y * ma.z;              // :1169-1170
        if (__float_as_int(ma.w) & 1) pn = 0.0f;                     // :1174-1176
        p[ca] = pn;
        bad = !isfinite(pn) || (check_max && fabsf(pn) > 2e3f);      // :1181-1192
    }
    if (vb) {
        float pn = pb * mb.x - bulk * db * mb.y * mb.z;
        if (__float_as_int(mb.w) & 1) pn = 0.0f;
        p[cb] = pn;
        bad = bad || !isfinite(pn) || (check_max && fabsf(pn) > 2e3f);
    }
    if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicMin(&status[0], s);
}

// Advance the device copy of the step count by the steps that completed.
__global__ void k_amr_step_count(int* status, int n) {
    status[2] += status[0] < n ? status[0] : n;
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

// Plate velocity boundary condition (acoustic_amr.cpp:1050-1072): each
// driven face takes +-(w - w_prev)/dt of the node that writes it last in the
// reference's node order (resolved on the host, one writer per face).
__global__ __launch_bounds__(256) void k_amr_plate_bc(const int* __restrict__ bc_face,
                                                      const int* __restrict__ bc_node,
                                                      const int* __restrict__ bc_sign, int n_bc,
                                                      const float* __restrict__ w,
                                                      const float* __restrict__ w_prev, float dt,
                                                      float* __restrict__ v,
                                                      const int* __restrict__ status, int s) {
    if (status[0] < s) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_bc) return;
    const int flat = bc_node[i];
    const float vp = (w[flat] - w_prev[flat]) / dt;
    v[bc_face[i]] = bc_sign[i] > 0 ? vp : -vp;
}

__device__ inline float plate_tap(const float* w, int i, int j, int Nx, int Ny) {
    return (i < 0 || i >= Nx || j < 0 || j >= Ny) ? 0.0f : w[i * Ny + j];
}

// 13-point biharmonic with zero outside the plate grid (acoustic_amr.cpp:939-954)
__device__ inline float plate_l4(const float* w, int i, int j, int Nx, int Ny, float idx4) {
    const float L = plate_tap(w, i - 2, j, Nx, Ny) + plate_tap(w, i + 2, j, Nx, Ny) +
                    plate_tap(w, i, j - 2, Nx, Ny) + plate_tap(w, i, j + 2, Nx, Ny) +
                    2.0f * (plate_tap(w, i - 1, j - 1, Nx, Ny) + plate_tap(w, i - 1, j + 1, Nx, Ny) +
                            plate_tap(w, i + 1, j - 1, Nx, Ny) + plate_tap(w, i + 1, j + 1, Nx, Ny)) -
                    8.0f * (plate_tap(w, i - 1, j, Nx, Ny) + plate_tap(w, i + 1, j, Nx, Ny) +
                            plate_tap(w, i, j - 1, Nx, Ny) + plate_tap(w, i, j + 1, Nx, Ny)) +
                    20.0f * plate_tap(w, i, j, Nx, Ny);
    return L * idx4;
}

struct PlateConsts {
    int Nx, Ny, n_active;
    float dx, dx4, dt2, inv_rh, damp_fwd, damp_bwd, cD0, cDp;
};

// One plate step (acoustic_amr.cpp:956-1047), one thread per active node;
// w_new is a separate buffer, so the rotation w_prev <- w <- w_new is a
// pointer swap on the host.  Each node clears its own external force.
__global__ __launch_bounds__(256) void k_amr_plate_step(
    PlateConsts k, const int* __restrict__ active, const int* __restrict__ ab_start,
    const int* __restrict__ ab_side, const float* __restrict__ ab_wgt,
    const int* __restrict__ bl_start, const int* __restrict__ bl_side,
    const float* __restrict__ bl_wgt, const int* __restrict__ c_above,
    const int* __restrict__ c_below, const float* __restrict__ p, const float* __restrict__ w,
    const float* __restrict__ w_prev, float* __restrict__ w_new, float* __restrict__ ext,
    const int* __restrict__ status, int s) {
    if (status[0] < s) return;
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= k.n_active) return;
    const int flat = active[n], i = flat / k.Ny, j = flat % k.Ny;
    float Pa = 0.0f, Pb = 0.0f;
    if (ab_start[n + 1] > ab_start[n]) {
        for (int q = ab_start[n]; q < ab_start[n + 1]; ++q) Pa += ab_wgt[q] * p[ab_side[q]];
    } else if (c_above[n] >= 0) {
        Pa = p[c_above[n]];
    }
    if (bl_start[n + 1] > bl_start[n]) {
        for (int q = bl_start[n]; q < bl_start[n + 1]; ++q) Pb += bl_wgt[q] * p[bl_side[q]];
    } else if (c_below[n] >= 0) {
        Pb = p[c_below[n]];
    }
    const float F_acou = (Pb - Pa) * k.dx * k.dx;
    const float F_ext = ext[flat];
    const float L4w = plate_l4(w, i, j, k.Nx, k.Ny, k.dx4);
    const float L4wp = plate_l4(w_prev, i, j, k.Nx, k.Ny, k.dx4);
    const float rhs = F_acou + F_ext - k.cD0 * L4w + k.cDp * L4wp;
    w_new[flat] = (k.damp_bwd * (2.0f * w[flat] - w_prev[flat]) + k.dt2 * k.inv_rh * rhs) /
                  k.damp_fwd;
    ext[flat] = 0.0f;
}

// External force from the bridge drive or a distributed force, accumulated
// in call order by one thread (entries may repeat a node).
__global__ void k_amr_plate_force(float* ext, const int* idx, const float* a, const float* b,
                                  float scale, int n, int bridge) {
    for (int q = 0; q < n; ++q)
        ext[idx[q]] += bridge ? a[q] * b[q] * scale : a[q] * scale;
}

__global__ __launch_bounds__(256) void k_amr_plate_sample4(const float* __restrict__ w,
                                                           const int* __restrict__ idx4,
                                                           const float* __restrict__ wgt4, int n,
                                                           float* __restrict__ out) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    float val = 0.0f;
    for (int k = 0; k < 4; ++k) val += wgt4[4 * q + k] * w[idx4[4 * q + k]];
    out[q] = val;
}

__global__ void k_amr_plate_weighted(const float* w, const int* idx, const float* wgt, int n,
                                     float* out) {
    float val = 0.0f;
    for (int q = 0; q < n; ++q) val += wgt[q] * w[idx[q]];
    out[0] = val;
}

// Mic samples (acoustic_amr.cpp:1762-1792): one 256-thread block per mic.
// The stencil's products are formed by the block's lanes in parallel, 256
// at a time, into LDS; lane 0 then adds them in stencil order, so each sum
// is the reference's sequential sum of the same rounded products (a single
// lane walking the stencil spent ~0.6 us per entry on the dependent loads).
struct MicSums {
    float p, vx, vy, vz;
};

__device__ MicSums mic_sums(int m, const int* __restrict__ mp_start,
                            const int* __restrict__ mp_cell, const float* __restrict__ mp_w,
                            const int* __restrict__ mv_start, const int* __restrict__ mv_face,
                            const float* __restrict__ wx, const float* __restrict__ wy,
                            const float* __restrict__ wz, const float* __restrict__ p,
                            const float* __restrict__ v) {
    __shared__ float sh[3][256];
    const int t = threadIdx.x;
    MicSums r{0.0f, 0.0f, 0.0f, 0.0f};
    const int p0 = mp_start[m], p1 = mp_start[m + 1];
    for (int base = p0; base < p1; base += 256) {
        const int k = base + t;
        if (k < p1) sh[0][t] = mp_w[k] * p[mp_cell[k]];
        __syncthreads();
        if (t == 0)
            for (int j = 0; j < min(256, p1 - base); ++j) r.p += sh[0][j];
        __syncthreads();
    }
    const int v0 = mv_start[m], v1 = mv_start[m + 1];
    for (int base = v0; base < v1; base += 256) {
        const int k = base + t;
        if (k < v1) {
            const float vf = v[mv_face[k]];
            sh[0][t] = wx[k] * vf;
            sh[1][t] = wy[k] * vf;
            sh[2][t] = wz[k] * vf;
        }
        __syncthreads();
        if (t == 0)
            for (int j = 0; j < min(256, v1 - base); ++j) {
                r.vx += sh[0][j];
                r.vy += sh[1][j];
                r.vz += sh[2][j];
            }
        __syncthreads();
    }
    return r;
}

__global__ __launch_bounds__(256) void k_amr_mics(
    const int* __restrict__ mp_start, const int* __restrict__ mp_cell,
    const float* __restrict__ mp_w, const int* __restrict__ mv_start,
    const int* __restrict__ mv_face, const float* __restrict__ wx, const float* __restrict__ wy,
    const float* __restrict__ wz, const float* __restrict__ p, const float* __restrict__ v,
    int n_mics, float* __restrict__ out) {
    const int m = blockIdx.x;
    const MicSums r = mic_sums(m, mp_start, mp_cell, mp_w, mv_start, mv_face, wx, wy, wz, p, v);
    if (threadIdx.x == 0) {
        out[m] = r.p;
        out[n_mics + m] = r.vx;
        out[2 * n_mics + m] = r.vy;
        out[3 * n_mics + m] = r.vz;
    }
}

// One audio sample of the co-evolver's mic output (acoustic_coevolver.cpp:
// 2208-2219): the mic samplers' p and v, mixed as
// (p*a + b*(rho c)*(v . axis)) * gain, into row `row`.  Skips when a step up
// to and including s_end failed (the co-evolver returns before sampling).
__global__ __launch_bounds__(256) void k_amr_mic_mix(
    const int* __restrict__ mp_start, const int* __restrict__ mp_cell,
    const float* __restrict__ mp_w, const int* __restrict__ mv_start,
    const int* __restrict__ mv_face, const float* __restrict__ wx, const float* __restrict__ wy,
    const float* __restrict__ wz, const float* __restrict__ p, const float* __restrict__ v,
    int n_mics, const float* __restrict__ mix, float rho_c, float* __restrict__ out, int row,
    const int* __restrict__ status, int s_end) {
    if (status[0] <= s_end) return;                  // uniform over the block
    const int m = blockIdx.x;
    const MicSums r = mic_sums(m, mp_start, mp_cell, mp_w, mv_start, mv_face, wx, wy, wz, p, v);
    if (threadIdx.x == 0) {
        const float* q = mix + 6 * m;                // a, b, axis x, y, z, gain
        float sig = r.p * q[0];
        if (q[1] != 0.0f) {
            const float vn = r.vx * q[2] + r.vy * q[3] + r.vz * q[4];
            sig += q[1] * rho_c * vn;
        }
        out[(size_t)row * n_mics + m] = sig * q[5];
    }
}

constexpr int kGraphSteps = 1024;

struct MicMix {                                      // per-sample mic recording
    int substeps;
    const float* mix;                                // device, 6 floats per mic
    float rho_c;
    float* out;                                      // device, n_samples x n_mics
};

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

// ── Gradient order 8: per-face Fornberg stencil (acoustic_amr.cpp:313-679) ──
// Built once on the host, as in the reference; each face's stencil depends
// only on the topology, so faces are split over host threads.

// The reference's first-derivative weight recurrence, as written
// (acoustic_amr.cpp:398-426: the new node's row uses x[n] - xi).
void fornberg_ref(const float* x, int N, float xi, float* w) {
    float c[16] = {};
    float c1 = 1.0f;
    c[0] = 1.0f;
    for (int n = 1; n < N; ++n) {
        const int mn = std::min(n, 1);
        float c2 = 1.0f;
        for (int nu = 0; nu < n; ++nu) {
            const float c3 = x[n] - x[nu];
            c2 *= c3;
            if (nu == n - 1) {
                for (int m = mn; m >= 1; --m)
                    c[2 * n + m] = c1 / c2 * ((float)m * c[2 * (n - 1) + m - 1] - (x[n] - xi) * c[2 * (n - 1) + m]);
                c[2 * n] = c1 / c2 * (-(x[n] - xi)) * c[2 * (n - 1)];
            }
            for (int m = mn; m >= 1; --m)
                c[2 * nu + m] = ((x[n] - xi) * c[2 * nu + m] - (float)m * c[2 * nu + m - 1]) / c3;
            c[2 * nu] = (x[n] - xi) * c[2 * nu] / c3;
        }
        c1 = c2;
    }
    for (int k = 0; k < N; ++k) w[k] = c[2 * k + 1];
}

struct StencilBuilder {
    const AcousticAMRState* st;
    const std::vector<int>& start;
    const std::vector<int>& csr;       // f, or ~f for the face's pos cell
    const std::vector<float>& ffc;
    std::vector<int> nb[6];            // +x, -x, +y, -y, +z, -z neighbour (-1: none)

    float proj(int ci, const float* d) const {
        const double* q = &st->centers[3 * (size_t)ci];
        return (float)(q[0] * d[0] + q[1] * d[1] + q[2] * d[2]);
    }
    float3 delta(int a, int b) const {    // centre b - centre a, per component in float
        const double* A = &st->centers[3 * (size_t)a];
        const double* B = &st->centers[3 * (size_t)b];
        return make_float3((float)(B[0] - A[0]), (float)(B[1] - A[1]), (float)(B[2] - A[2]));
    }
    static int axis_of(float ax, float ay, float az) {
        const float best = std::max({ax, ay, az});
        if (!(best > 0.9f)) return -1;
        if (ax >= ay && ax >= az) return 0;
        if (ay >= ax && ay >= az) return 1;
        return 2;
    }
    bool blocked(int c) const { return st->type[c] == 1 || st->type[c] == 2; }

    // acoustic_amr.cpp:329-391
    void build_neighbours() {
        std::vector<uint8_t> state[6];
        for (int q = 0; q < 6; ++q) {
            nb[q].assign(st->nc, -1);
            state[q].assign(st->nc, 0);
        }
        auto set = [&](int q, int c, int val) {
            int& t = nb[q][c];
            uint8_t& k = state[q][c];
            if (k == 0) { t = val; k = 1; }
            else if (k == 1 && t != val) { t = -1; k = 2; }
        };
        for (int f = 0; f < st->nf; ++f) {
            const int cn = st->neg[f], cp = st->pos[f];
            const float3 d = delta(cn, cp);
            const float len = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
            if (len < 1e-15f) continue;
            const int ax = axis_of(fabsf(d.x / len), fabsf(d.y / len), fabsf(d.z / len));
            if (ax < 0) continue;
            const float sd = ax == 0 ? d.x : (ax == 1 ? d.y : d.z);
            const int lo = sd > 0.0f ? cn : cp, hi = sd > 0.0f ? cp : cn;
            set(2 * ax, lo, hi);
            set(2 * ax + 1, hi, lo);
        }
    }

    // acoustic_amr.cpp:444-547
    int walk(int seed, float fproj, const float* cd, const float* wd, int* oc, float* ox) const {
        oc[0] = seed;
        ox[0] = proj(seed, cd) - fproj;
        int found = 1, cur = seed;
        const int ax = axis_of(fabsf(wd[0]), fabsf(wd[1]), fabsf(wd[2]));
        if (ax >= 0) {
            const std::vector<int>& t = nb[2 * ax + (wd[ax] > 0.0f ? 0 : 1)];
            while (found < 4) {
                const int n = t[cur];
                if (n < 0 || n >= st->nc || blocked(n)) break;
                oc[found] = n;
                ox[found] = proj(n, cd) - fproj;
                ++found;
                cur = n;
            }
            return found;
        }
        while (found < 4) {
            int best_next = -1;
            float best_dot = 0.5f;
            for (int k = start[cur]; k < start[cur + 1]; ++k) {
                const int e = csr[k];
                const int fk = e < 0 ? ~e : e;
                const float wk = e < 0 ? -ffc[fk] : ffc[fk];  // the reference's signed weight
                const float sg = wk > 0.f ? 1.f : -1.f;
                const int fn = st->neg[fk], fp = st->pos[fk];
                const float3 d = delta(fn, fp);
                const float len = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
                if (len < 1e-15f) continue;
                const float dot = sg * (d.x * wd[0] + d.y * wd[1] + d.z * wd[2]) / len;
                if (dot > best_dot) {
                    const int n = sg > 0.f ? fp : fn;
                    if (n >= 0 && n < st->nc && !blocked(n)) {
                        best_dot = dot;
                        best_next = n;
                    }
                }
            }
            if (best_next < 0) break;
            oc[found] = best_next;
            ox[found] = proj(best_next, cd) - fproj;
            ++found;
            cur = best_next;
        }
        return found;
    }

    // acoustic_amr.cpp:549-629, for faces [f0, f1)
    void faces(int f0, int f1, int* sc, float* sw) const {
        for (int f = f0; f < f1; ++f) {
            const int cn = st->neg[f], cp = st->pos[f];
            const float3 d = delta(cn, cp);
            const float len = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
            if (!(len >= 1e-15f)) continue;
            const float dir[3] = {d.x / len, d.y / len, d.z / len};
            const float mdir[3] = {-dir[0], -dir[1], -dir[2]};
            const float fproj = 0.5f * (proj(cn, dir) + proj(cp, dir));
            int nc4[4], pc4[4];
            float nx4[4], px4[4];
            const int nn = walk(cn, fproj, dir, mdir, nc4, nx4);
            const int np = walk(cp, fproj, dir, dir, pc4, px4);
            int* c8 = sc + 8 * (size_t)f;
            float* w8 = sw + 8 * (size_t)f;
            if (nn == 4 && np == 4) {
                float xs[8], w[8];
                for (int k = 0; k < 4; ++k) {
                    xs[k] = nx4[3 - k];
                    xs[4 + k] = px4[k];
                }
                fornberg_ref(xs, 8, 0.0f, w);
                for (int k = 0; k < 4; ++k) {
                    c8[k] = nc4[k];
                    w8[k] = w[3 - k];
                    c8[4 + k] = pc4[k];
                    w8[4 + k] = w[4 + k];
                }
            } else if (nn + np >= 2) {
                c8[0] = cn;
                w8[0] = -st->inv_d[f];
                c8[4] = cp;
                w8[4] = st->inv_d[f];
            }
        }
    }
};

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
    void* comp[] = {st->d_fc, st->d_fcls, st->d_cc, st->d_ccls};
    for (void* b : comp)
        if (b) (void)hipFree(b);
    void* more[] = {st->d_sc, st->d_sw, st->d_pw[0], st->d_pw[1], st->d_pw[2], st->d_ext, st->d_pint,
                    st->d_pflt, st->d_mint, st->d_mflt, st->d_mout, st->d_tmp};
    for (void* b : more)
        if (b) (void)hipFree(b);
    drop_graph(st);
    if (st->stream) (void)hipStreamDestroy(st->stream);
    delete st;
}

// Distinct-record tables.  Keys are the records' exact bit patterns, so a
// class holds exactly the values the direct record would.
template <int K>
struct KeyHash {
    size_t operator()(const std::array<uint32_t, K>& k) const {
        uint64_t h = 1469598103934665603ull;
        for (uint32_t x : k) h = (h ^ x) * 1099511628211ull;
        return (size_t)(h ^ (h >> 29));
    }
};

template <int K>
struct ClassTable {
    std::unordered_map<std::array<uint32_t, K>, int, KeyHash<K>> ids;
    std::vector<std::array<uint32_t, K>> rows;
    int limit;
    explicit ClassTable(int lim) : limit(lim) {}
    int id(const std::array<uint32_t, K>& k) {       // -1 once over the limit
        auto it = ids.find(k);
        if (it != ids.end()) return it->second;
        if ((int)rows.size() >= limit) return -1;
        ids.emplace(k, (int)rows.size());
        rows.push_back(k);
        return (int)rows.size() - 1;
    }
};

constexpr int kMaxClasses = 1 << 16;
constexpr uint32_t kAbsZero = 0x80000000u;   // order-8 slot that names cell 0 itself

inline uint32_t bits(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    return u;
}

template <class T>
hipError_t upload_vec(T** dst, const std::vector<T>& v, hipStream_t s) {
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    hipError_t e = hipMalloc((void**)dst, sizeof(T) * (v.size() + 1));
    if (e == hipSuccess && !v.empty())
        e = hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s);
    return e;
}

// Build the compact records from the direct ones; falls back to the direct
// layout when requested or when a table would exceed kMaxClasses.
hipError_t build_compact(AcousticAMRState* st, const std::vector<int4>& face,
                         const std::vector<float4>& cell) {
    st->compact = false;
    if (st->layout_req == 1) return hipSuccess;
    const int lim = st->layout_req == 2 ? (1 << 24) : kMaxClasses;
    std::vector<int2> fc(st->nf);
    std::vector<int4> fcls;
    bool ok = true;
    if (st->order == 8) {
        ClassTable<17> t(lim);
        for (int f = 0; f < st->nf && ok; ++f) {
            std::array<uint32_t, 17> k;
            for (int q = 0; q < 8; ++q) {
                const int c = st->sc8[8 * (size_t)f + q];
                k[q] = c == 0 ? kAbsZero : (uint32_t)(c - st->neg[f]);
                k[8 + q] = bits(st->sw8[8 * (size_t)f + q]);
            }
            k[16] = (uint32_t)face[f].w;
            const int id = t.id(k);
            ok = id >= 0;
            fc[f] = make_int2(st->neg[f], id);
        }
        if (ok)
            for (auto& r : t.rows) {
                fcls.push_back(make_int4(r[0], r[1], r[2], r[3]));
                fcls.push_back(make_int4(r[4], r[5], r[6], r[7]));
                fcls.push_back(make_int4(r[8], r[9], r[10], r[11]));
                fcls.push_back(make_int4(r[12], r[13], r[14], r[15]));
                fcls.push_back(make_int4(r[16], 0, 0, 0));
            }
        st->n_fclass = (int)t.rows.size();
    } else {
        ClassTable<3> t(lim);
        for (int f = 0; f < st->nf && ok; ++f) {
            const int4 r = face[f];
            const int id = t.id({(uint32_t)(r.y - r.x), (uint32_t)r.z, (uint32_t)r.w});
            ok = id >= 0;
            fc[f] = make_int2(r.x, id);
        }
        if (ok)
            for (auto& r : t.rows) fcls.push_back(make_int4(r[0], r[1], r[2], 0));
        st->n_fclass = (int)t.rows.size();
    }
    std::vector<int> cc(st->nc);
    std::vector<float4> ccls;
    if (ok) {
        ClassTable<4> t(lim);
        for (int i = 0; i < st->nc && ok; ++i) {
            const float4 m = cell[i];
            cc[i] = t.id({bits(m.x), bits(m.y), bits(m.z), bits(m.w)});
            ok = cc[i] >= 0;
        }
        if (ok)
            for (auto& r : t.rows) {
                float4 m;
                memcpy(&m, r.data(), 16);
                ccls.push_back(m);
            }
        st->n_cclass = (int)t.rows.size();
    }
    if (!ok) return hipSuccess;                      // too many distinct records: direct
    hipError_t e = upload_vec(&st->d_fc, fc, st->stream);
    if (e == hipSuccess) e = upload_vec(&st->d_fcls, fcls, st->stream);
    if (e == hipSuccess) e = upload_vec(&st->d_cc, cc, st->stream);
    if (e == hipSuccess) e = upload_vec(&st->d_ccls, ccls, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    st->compact = e == hipSuccess;
    return e;
}

// Pack the per-cell and per-face coefficient records and upload them.
hipError_t upload_coeffs(AcousticAMRState* st) {
    drop_graph(st);
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
    if (e == hipSuccess) e = build_compact(st, face, cell);
    ++st->gen;
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

// Enqueue the kernels of step s of a call (s counts from the call's first
// step; each kernel skips when an earlier step of the call failed): faces,
// plate-driven faces, cells, plate.  `mid` (optional) is recorded between
// the face and the cell kernels.
void enqueue_amr_step(AcousticAMRState* st, int s, float dt, float bulk, float dt_over_rho,
                      hipEvent_t mid) {
    if (st->order == 8 && st->compact)
        k_amr_faces8_c<<<(st->nf + kFaceTile8C - 1) / kFaceTile8C, 256, 0, st->stream>>>(
            st->d_fc, st->d_fcls, st->d_v, st->d_p, st->nf, dt_over_rho, st->d_status, s);
    else if (st->order == 8)
        k_amr_faces8<<<blocks(st->nf), 256, 0, st->stream>>>(
            st->d_sc, st->d_sw, st->d_face, st->d_v, st->d_p, st->nf, dt_over_rho,
            st->d_status, s);
    else if (st->compact)
        k_amr_faces_c<<<(st->nf + kFaceTileC - 1) / kFaceTileC, 256, 0, st->stream>>>(
            st->d_fc, st->d_fcls, st->d_v, st->d_p, st->nf, dt_over_rho, st->d_status, s);
    else
        k_amr_faces<<<(st->nf + kFaceTileD - 1) / kFaceTileD, 256, 0, st->stream>>>(
            st->d_face, st->d_v, st->d_p, st->nf, dt_over_rho, st->d_status, s);
    if (st->plate_on && st->n_bc > 0) {                      // :1157
        const float* w = st->d_pw[(st->prot + s + 1) % 3];
        const float* wp = st->d_pw[(st->prot + s) % 3];
        k_amr_plate_bc<<<blocks(st->n_bc), 256, 0, st->stream>>>(
            st->d_pint + st->o_bcf, st->d_pint + st->o_bcn, st->d_pint + st->o_bcs,
            st->n_bc, w, wp, dt, st->d_v, st->d_status, s);
    }
    if (mid) (void)hipEventRecord(mid, st->stream);
    const unsigned cell_blocks = (unsigned)((st->nc + kCellTile - 1) / kCellTile);
    if (st->compact)
        k_amr_cells<true><<<cell_blocks, 256, 0, st->stream>>>(
            st->d_ccls, st->d_cc, st->d_start, st->d_csr, st->d_w, st->d_v, st->d_p, st->nc,
            bulk, st->d_status, s);
    else
        k_amr_cells<false><<<cell_blocks, 256, 0, st->stream>>>(
            st->d_cell, nullptr, st->d_start, st->d_csr, st->d_w, st->d_v, st->d_p, st->nc,
            bulk, st->d_status, s);
    if (st->plate_on && st->n_active > 0) {                  // :1179
        PlateConsts k;
        k.Nx = st->pNx;
        k.Ny = st->pNy;
        k.n_active = st->n_active;
        k.dx = st->p_dx;
        k.dx4 = 1.0f / (st->p_dx * st->p_dx * st->p_dx * st->p_dx);
        k.dt2 = dt * dt;
        k.inv_rh = 1.0f / st->p_rh;
        k.damp_fwd = 1.0f + 0.5f * st->p_aM * dt;
        k.damp_bwd = 1.0f - 0.5f * st->p_aM * dt;
        k.cD0 = st->p_D * (1.0f + st->p_bK / dt);
        k.cDp = st->p_D * (st->p_bK / dt);
        const int* I = st->d_pint;
        const float* F = st->d_pflt;
        k_amr_plate_step<<<blocks(st->n_active), 256, 0, st->stream>>>(
            k, I + st->o_active, I + st->o_ab, I + st->o_abs, F + st->o_abw, I + st->o_bl,
            I + st->o_bls, F + st->o_blw, I + st->o_ca, I + st->o_cb, st->d_p,
            st->d_pw[(st->prot + s + 1) % 3], st->d_pw[(st->prot + s) % 3],
            st->d_pw[(st->prot + s + 2) % 3], st->d_ext, st->d_status, s);
    }
}

int step_impl(AcousticAMRState* st, int n, float* ms_total, float* ms_faces, float* ms_cells,
              const MicMix* mix = nullptr) {
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
    // A block of recorded samples of at most kGraphSteps steps is replayed
    // from a captured graph while the same block shape runs again.  Measured
    // gain is small (1-5 % on 33 K - 885 K-cell grids with 64-sample
    // blocks: the floor is each kernel's own latency, not its launch), and
    // capturing a long block costs more than it saves, hence the limit.
    bool graph = false, replay = false;
    std::array<uint64_t, 6> gkey{};
    const int slot = (st->plate_on && st->n_active > 0) ? st->prot % 3 : 0;
    if (mix && !ms_total && n <= kGraphSteps) {
        gkey = {(uint64_t)n, (uint64_t)mix->substeps, (uint64_t)slot, st->gen,
                (uint64_t)(uintptr_t)mix->out, (uint64_t)bits(mix->rho_c)};
        replay = true;
        if (!st->gexec[slot] || st->gkey[slot] != gkey) {
            if (st->gexec[slot]) (void)hipGraphExecDestroy(st->gexec[slot]);
            st->gexec[slot] = nullptr;
            if (hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return AMR_ERR_DEVICE;
            graph = true;
        }
    }
    if (graph || !replay) {
    k_amr_status_init<<<1, 1, 0, st->stream>>>(st->d_status, st->d_best);
    for (int s = 0; s < n; ++s) {
        if (ms_total) (void)hipEventRecord(ev[2 * (size_t)s], st->stream);
        enqueue_amr_step(st, s, dt, bulk, dt_over_rho, ms_total ? ev[2 * (size_t)s + 1] : nullptr);
        if (mix && (s + 1) % mix->substeps == 0) {
            const int* I = st->d_mint;
            const float* F = st->d_mflt;
            k_amr_mic_mix<<<st->n_mics, 256, 0, st->stream>>>(
                I, I + st->o_mpc, F, I + st->o_mvs, I + st->o_mvf, F + st->o_mwx, F + st->o_mwy,
                F + st->o_mwz, st->d_p, st->d_v, st->n_mics, mix->mix, mix->rho_c, mix->out,
                (s + 1) / mix->substeps - 1, st->d_status, s);
        }
    }
    k_amr_step_count<<<1, 1, 0, st->stream>>>(st->d_status, n);
    }
    if (graph) {
        hipGraph_t gr = nullptr;
        hipError_t ge = hipStreamEndCapture(st->stream, &gr);
        if (ge == hipSuccess) ge = hipGraphInstantiate(&st->gexec[slot], gr, nullptr, nullptr, 0);
        if (gr) (void)hipGraphDestroy(gr);
        if (ge != hipSuccess) {
            st->gexec[slot] = nullptr;
            return AMR_ERR_DEVICE;
        }
        st->gkey[slot] = gkey;
    }
    if (replay && hipGraphLaunch(st->gexec[slot], st->stream) != hipSuccess) return AMR_ERR_DEVICE;
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
    // plate buffers rotate once per step whose plate update ran (the failing
    // step's included, as the reference advances the plate before its test)
    if (st->plate_on && st->n_active > 0) st->prot = (st->prot + (fail < n ? fail + 1 : n)) % 3;
    if (fail < n) {                                  // steps before it completed
        st->step_count += fail;
        report_diverged(st);
        return SK_ERR_DIVERGED;
    }
    st->step_count += n;
    return SK_OK;
}

}  // namespace

__global__ void k_amr_hook_init(int* status) { status[3] = INT_MAX; }

// Device step count after a hooked call: the steps before a failed hook, or
// before a diverged step, or all n.
__global__ void k_amr_step_count_hooked(int* status, int n) {
    status[2] += status[3] != INT_MAX ? status[3] : (status[0] < n ? status[0] : n);
}

int amr_step_hooked(AcousticAMRState* st, int n, const AmrStepHook& hook, float* ms_total,
                    int* done, int* hook_failed) {
    *done = 0;
    *hook_failed = 0;
    if (ms_total) *ms_total = 0.0f;
    if (!st) return SK_ERR_NULL_STATE;
    if (n < 0) return SK_ERR_DIM_MISMATCH;
    const float dt = (float)st->dt;
    const float bulk = (float)(st->rho * st->c * st->c * st->dt);
    const float dt_over_rho = dt / (float)st->rho;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms_total && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        return AMR_ERR_DEVICE;
    auto args = [&](int q) {
        AmrHookArgs a;
        a.q = q;
        a.w = st->plate_on ? st->d_pw[(st->prot + q + 1) % 3] : nullptr;
        a.ext = st->d_ext;
        a.p = st->d_p;
        a.v = st->d_v;
        return a;
    };
    int rc = SK_OK;
    k_amr_status_init<<<1, 1, 0, st->stream>>>(st->d_status, st->d_best);
    k_amr_hook_init<<<1, 1, 0, st->stream>>>(st->d_status);
    if (e0) (void)hipEventRecord(e0, st->stream);
    for (int s = 0; s < n && rc == SK_OK; ++s) {
        rc = hook(args(s));
        if (rc == SK_OK) enqueue_amr_step(st, s, dt, bulk, dt_over_rho, nullptr);
    }
    if (rc == SK_OK) rc = hook(args(n));
    k_amr_step_count_hooked<<<1, 1, 0, st->stream>>>(st->d_status, n);
    if (e1) (void)hipEventRecord(e1, st->stream);
    int status[4] = {INT_MAX, INT_MAX, 0, INT_MAX};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(status, st->d_status, sizeof(status), hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    if (e == hipSuccess && ms_total) (void)hipEventElapsedTime(ms_total, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess || rc != SK_OK) return AMR_ERR_DEVICE;
    const bool rot = st->plate_on && st->n_active > 0;
    if (status[3] != INT_MAX) {                      // the hook failed before step status[3]
        *done = status[3];
        *hook_failed = 1;
        st->step_count += status[3];
        if (rot) st->prot = (st->prot + status[3]) % 3;
        return SK_ERR_DIVERGED;
    }
    if (status[0] < n) {                             // step status[0] diverged
        *done = status[0];
        st->step_count += status[0];
        if (rot) st->prot = (st->prot + status[0] + 1) % 3;
        report_diverged(st);
        return SK_ERR_DIVERGED;
    }
    *done = n;
    st->step_count += n;
    if (rot) st->prot = (st->prot + n) % 3;
    return SK_OK;
}

int amr_co_view(AcousticAMRState* st, AmrCoView* v) {
    if (!st || !v) return SK_ERR_NULL_STATE;
    v->stream = st->stream;
    v->status = st->d_status;
    v->n_cells = st->nc;
    v->n_pnodes = st->n_pnodes;
    v->pNx = st->pNx;
    v->pNy = st->pNy;
    v->n_mics = st->n_mics;
    const int* I = st->d_mint;
    const float* F = st->d_mflt;
    const bool m = st->n_mics > 0 && I && F;
    v->n_mp = m ? st->o_mvf - st->o_mpc : 0;
    v->n_mv = m ? st->o_mwy - st->o_mwx : 0;
    v->mp_start = m ? I : nullptr;
    v->mp_cell = m ? I + st->o_mpc : nullptr;
    v->mp_w = m ? F : nullptr;
    v->mv_start = m ? I + st->o_mvs : nullptr;
    v->mv_face = m ? I + st->o_mvf : nullptr;
    v->mv_w[0] = m ? F + st->o_mwx : nullptr;
    v->mv_w[1] = m ? F + st->o_mwy : nullptr;
    v->mv_w[2] = m ? F + st->o_mwz : nullptr;
    return SK_OK;
}

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
    auto* st = new AcousticAMRState();
    st->nc = nc;
    st->nf = nf;
    st->order = order;
    st->c = c;
    st->rho = rho;
    st->dt = (order == 8 ? 0.10 : 0.77) * min_dx / (c * sqrt(3.0));  // :744-745
    st->centers.assign(cc, cc + 3 * (size_t)nc);
    st->open_vol.assign(ov, ov + nc);
    st->type.assign(ct, ct + nc);
    st->neg.assign(fn, fn + nf);
    st->pos.assign(fp, fp + nf);
    st->volumes.assign(vol, vol + nc);
    st->face_area.assign(fa, fa + nf);
    st->face_open.assign(fo, fo + nf);
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
    std::vector<int>& sc8 = st->sc8;
    std::vector<float>& sw8 = st->sw8;
    if (order == 8) {                                                // :862-869
        set_stage("stencil.build", "");
        sc8.assign(8 * (size_t)nf, 0);
        sw8.assign(8 * (size_t)nf, 0.0f);
        StencilBuilder sb{st, start, csr, ffc, {}};
        sb.build_neighbours();
        g_total_faces = nf;
        const int nt = std::max(1, std::min(64, (int)std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (int t = 0; t < (nf < 256 ? 1 : nt); ++t) {
            const int n = nf < 256 ? 1 : nt;
            const int f0 = (int)((int64_t)nf * t / n), f1 = (int)((int64_t)nf * (t + 1) / n);
            pool.emplace_back([&sb, f0, f1, &sc8, &sw8] { sb.faces(f0, f1, sc8.data(), sw8.data()); });
        }
        for (auto& th : pool) th.join();
        g_done_faces = nf;
        set_stage("ok", "");
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
    if (order == 8) {
        al((void**)&st->d_sc, sizeof(int) * 8 * (size_t)nf);
        al((void**)&st->d_sw, sizeof(float) * 8 * (size_t)nf);
    }
    al((void**)&st->d_status, sizeof(int) * 4);      // fail step, nearest cell, step count
    al((void**)&st->d_best, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemsetAsync(st->d_v, 0, sizeof(float) * (size_t)nf, st->stream);
    if (e == hipSuccess) e = hipMemsetAsync(st->d_p, 0, sizeof(float) * (size_t)nc, st->stream);
    if (e == hipSuccess) e = hipMemsetAsync(st->d_status, 0, sizeof(int) * 4, st->stream);
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
    if (e == hipSuccess && order == 8)
        e = hipMemcpyAsync(st->d_sc, sc8.data(), sizeof(int) * 8 * (size_t)nf,
                           hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess && order == 8)
        e = hipMemcpyAsync(st->d_sw, sw8.data(), sizeof(float) * 8 * (size_t)nf,
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
SK_API int amr_get_create_progress_active(void) { return 0; }   // the build is synchronous
SK_API int amr_get_create_progress_done_faces(void) { return g_done_faces.load(); }
SK_API int amr_get_create_progress_total_faces(void) { return g_total_faces.load(); }

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
    if (e == hipSuccess) e = hipMemsetAsync(st->d_status, 0, sizeof(int) * 4, st->stream);
    if (st->plate_on) {
        for (int b = 0; b < 3 && e == hipSuccess; ++b)
            e = hipMemsetAsync(st->d_pw[b], 0, sizeof(float) * (size_t)st->n_pnodes, st->stream);
        if (e == hipSuccess)
            e = hipMemsetAsync(st->d_ext, 0, sizeof(float) * (size_t)st->n_pnodes, st->stream);
    }
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

SK_API int amr_set_record_layout(AcousticAMRState* st, int layout) {
    if (!st) return SK_ERR_NULL_STATE;
    if (layout < 0 || layout > 2) return SK_ERR_DIM_MISMATCH;
    st->layout_req = layout;
    return upload_coeffs(st) == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

SK_API int amr_get_record_layout(const AcousticAMRState* st) {
    return st ? (st->compact ? 2 : 1) : 0;
}

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

// ── Kirchhoff plate (acoustic_amr.cpp:1412-1644) ────────────────────────

static hipError_t ensure_tmp(AcousticAMRState* st, size_t bytes) {
    if (bytes <= st->tmp_bytes) return hipSuccess;
    drop_graph(st);                 // captured mic blocks reference d_tmp
    if (st->d_tmp) (void)hipFree(st->d_tmp);
    st->d_tmp = nullptr;
    st->tmp_bytes = 0;
    hipError_t e = hipMalloc(&st->d_tmp, bytes);
    if (e == hipSuccess) st->tmp_bytes = bytes;
    ++st->gen;
    return e;
}

SK_API int amr_setup_plate(AcousticAMRState* st, int Nx, int Ny, float pdx, const float origin[3],
                           const uint8_t* act, float rh, float D, float aM, float bK,
                           int n_active, const int32_t* flat_idx, const int32_t* ab_start,
                           const int32_t* ab_idx, const float* ab_w, const int32_t* bl_start,
                           const int32_t* bl_idx, const float* bl_w, const int32_t* c_above,
                           const int32_t* c_below) {
    (void)origin;
    if (!st) return SK_ERR_NULL_STATE;
    if (!act || !flat_idx || !ab_start || !bl_start || !c_above || !c_below)
        return SK_ERR_NULL_STATE;
    if (Nx <= 0 || Ny <= 0 || pdx <= 0.0f) return SK_ERR_DIM_MISMATCH;
    if (rh <= 0.0f || D <= 0.0f) return SK_ERR_DIM_MISMATCH;
    if (n_active < 0) return SK_ERR_DIM_MISMATCH;
    const int nn = Nx * Ny;
    int cnt = 0;
    for (int k = 0; k < nn; ++k) cnt += act[k] ? 1 : 0;
    if (cnt != n_active) return SK_ERR_DIM_MISMATCH;
    for (int k = 0; k < ab_start[n_active]; ++k)
        if (ab_idx[k] < 0 || ab_idx[k] >= st->nf) return SK_ERR_DIM_MISMATCH;
    for (int k = 0; k < bl_start[n_active]; ++k)
        if (bl_idx[k] < 0 || bl_idx[k] >= st->nf) return SK_ERR_DIM_MISMATCH;
    for (int n = 0; n < n_active; ++n) {
        if (c_above[n] != -1 && (c_above[n] < 0 || c_above[n] >= st->nc)) return SK_ERR_DIM_MISMATCH;
        if (c_below[n] != -1 && (c_below[n] < 0 || c_below[n] >= st->nc)) return SK_ERR_DIM_MISMATCH;
        // not checked by the reference; needed to keep device accesses in bounds
        if (flat_idx[n] < 0 || flat_idx[n] >= nn) return SK_ERR_DIM_MISMATCH;
    }
    // node maps, face lists rebased to start at 0, air cell of each face
    const int a0 = ab_start[0], b0 = bl_start[0];
    const int la = ab_start[n_active] - a0, lb = bl_start[n_active] - b0;
    std::vector<int> I;
    auto put = [&](int& off, const int* src, int n) {
        off = (int)I.size();
        I.insert(I.end(), src, src + n);
    };
    std::vector<int> abs_(n_active + 1), bls_(n_active + 1), aside(la), bside(lb);
    for (int n = 0; n <= n_active; ++n) { abs_[n] = ab_start[n] - a0; bls_[n] = bl_start[n] - b0; }
    for (int k = 0; k < la; ++k) aside[k] = st->pos[ab_idx[a0 + k]];
    for (int k = 0; k < lb; ++k) bside[k] = st->neg[bl_idx[b0 + k]];
    // driven faces: the last writer in node order (above faces, then below) wins
    std::vector<int> owner_node(st->nf, -1), owner_sign(st->nf, 0), order;
    for (int n = 0; n < n_active; ++n) {
        for (int k = abs_[n]; k < abs_[n + 1]; ++k) {
            const int f = ab_idx[a0 + k];
            if (owner_node[f] < 0) order.push_back(f);
            owner_node[f] = flat_idx[n];
            owner_sign[f] = 1;
        }
        for (int k = bls_[n]; k < bls_[n + 1]; ++k) {
            const int f = bl_idx[b0 + k];
            if (owner_node[f] < 0) order.push_back(f);
            owner_node[f] = flat_idx[n];
            owner_sign[f] = -1;
        }
    }
    std::vector<int> bcn(order.size()), bcs(order.size());
    for (size_t q = 0; q < order.size(); ++q) {
        bcn[q] = owner_node[order[q]];
        bcs[q] = owner_sign[order[q]];
    }
    int o_active, o_ab, o_bl, o_abs, o_bls, o_ca, o_cb, o_bcf, o_bcn, o_bcs;
    put(o_active, flat_idx, n_active);
    put(o_ab, abs_.data(), n_active + 1);
    put(o_bl, bls_.data(), n_active + 1);
    put(o_abs, aside.data(), la);
    put(o_bls, bside.data(), lb);
    put(o_ca, c_above, n_active);
    put(o_cb, c_below, n_active);
    put(o_bcf, order.data(), (int)order.size());
    put(o_bcn, bcn.data(), (int)bcn.size());
    put(o_bcs, bcs.data(), (int)bcs.size());
    std::vector<float> F(ab_w + a0, ab_w + a0 + la);
    F.insert(F.end(), bl_w + b0, bl_w + b0 + lb);

    // replace any previous plate
    drop_graph(st);
    for (int b = 0; b < 3; ++b) {
        if (st->d_pw[b]) (void)hipFree(st->d_pw[b]);
        st->d_pw[b] = nullptr;
    }
    if (st->d_ext) (void)hipFree(st->d_ext);
    if (st->d_pint) (void)hipFree(st->d_pint);
    if (st->d_pflt) (void)hipFree(st->d_pflt);
    st->d_ext = nullptr;
    st->d_pint = nullptr;
    st->d_pflt = nullptr;
    const bool was_on = st->plate_on;
    st->plate_on = false;
    hipError_t e = hipSuccess;
    for (int b = 0; b < 3 && e == hipSuccess; ++b) e = hipMalloc((void**)&st->d_pw[b], sizeof(float) * nn);
    if (e == hipSuccess) e = hipMalloc((void**)&st->d_ext, sizeof(float) * nn);
    if (e == hipSuccess) e = hipMalloc((void**)&st->d_pint, sizeof(int) * (I.size() + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&st->d_pflt, sizeof(float) * (F.size() + 1));
    for (int b = 0; b < 3 && e == hipSuccess; ++b)
        e = hipMemsetAsync(st->d_pw[b], 0, sizeof(float) * nn, st->stream);
    if (e == hipSuccess) e = hipMemsetAsync(st->d_ext, 0, sizeof(float) * nn, st->stream);
    if (e == hipSuccess && !I.empty())
        e = hipMemcpyAsync(st->d_pint, I.data(), sizeof(int) * I.size(), hipMemcpyHostToDevice,
                           st->stream);
    if (e == hipSuccess && !F.empty())
        e = hipMemcpyAsync(st->d_pflt, F.data(), sizeof(float) * F.size(), hipMemcpyHostToDevice,
                           st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    if (e != hipSuccess) return AMR_ERR_DEVICE;
    st->pNx = Nx;
    st->pNy = Ny;
    st->n_pnodes = nn;
    st->n_active = n_active;
    st->p_dx = pdx;
    st->p_rh = rh;
    st->p_D = D;
    st->p_aM = aM;
    st->p_bK = bK;
    st->prot = 0;
    st->n_bc = (int)order.size();
    st->o_active = o_active; st->o_ab = o_ab; st->o_bl = o_bl; st->o_abs = o_abs;
    st->o_bls = o_bls; st->o_ca = o_ca; st->o_cb = o_cb; st->o_bcf = o_bcf;
    st->o_bcn = o_bcn; st->o_bcs = o_bcs;
    st->o_abw = 0;
    st->o_blw = la;
    ++st->gen;
    for (int n = 0; n < n_active; ++n)               // every node coupled on both sides
        if ((abs_[n + 1] == abs_[n] && c_above[n] < 0) || (bls_[n + 1] == bls_[n] && c_below[n] < 0)) {
            st->plate_on = was_on;                   // the reference leaves its flag as it was
            return SK_ERR_DIM_MISMATCH;
        }
    st->plate_on = true;
    return SK_OK;
}

SK_API int amr_set_bridge_plate_sources(AcousticAMRState* st, int n, const int32_t* idx,
                                        const float* wgt) {
    if (!st) return SK_ERR_NULL_STATE;                               // :1534-1546
    if (!st->plate_on) return SK_ERR_DIM_MISMATCH;
    if (n <= 0 || !idx || !wgt) return SK_ERR_DIM_MISMATCH;
    for (int i = 0; i < n; ++i)
        if (idx[i] < 0 || idx[i] >= st->n_pnodes) return SK_ERR_DIM_MISMATCH;
    st->bridge_idx.assign(idx, idx + n);
    st->bridge_wgt.assign(wgt, wgt + n);
    return SK_OK;
}

SK_API int amr_set_neck_plate_sources(AcousticAMRState* st, int n, const int32_t* idx,
                                      const float* wgt) {
    if (!st) return SK_ERR_NULL_STATE;                               // :1548-1563
    if (n == 0) {
        st->neck_idx.clear();
        st->neck_wgt.clear();
        return SK_OK;
    }
    if (!st->plate_on) return SK_ERR_DIM_MISMATCH;
    if (!idx || !wgt) return SK_ERR_DIM_MISMATCH;
    for (int i = 0; i < n; ++i)
        if (idx[i] < 0 || idx[i] >= st->n_pnodes) return SK_ERR_DIM_MISMATCH;
    st->neck_idx.assign(idx, idx + n);
    st->neck_wgt.assign(wgt, wgt + n);
    return SK_OK;
}

// Upload n (index, a, b) triples and accumulate them into the external force.
static int plate_force(AcousticAMRState* st, const int* idx, const float* a, const float* b,
                       float scale, int n, int bridge) {
    if (n <= 0) return SK_OK;
    const size_t bytes = (size_t)n * (sizeof(int) + 2 * sizeof(float));
    hipError_t e = ensure_tmp(st, bytes);
    int* di = (int*)st->d_tmp;
    float* da = (float*)(di + n);
    float* db = da + n;
    if (e == hipSuccess)
        e = hipMemcpyAsync(di, idx, sizeof(int) * n, hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(da, a, sizeof(float) * n, hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess && b)
        e = hipMemcpyAsync(db, b, sizeof(float) * n, hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess) {
        k_amr_plate_force<<<1, 1, 0, st->stream>>>(st->d_ext, di, da, db, scale, n, bridge);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);      // caller may reuse buffers
    return e == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

SK_API int amr_inject_bridge_drive(AcousticAMRState* st, const float* drives, int n, float scale) {
    if (!st || !drives) return SK_ERR_NULL_STATE;                    // :1567-1579
    if (!st->plate_on) return SK_ERR_DIM_MISMATCH;
    if (n != (int)st->bridge_idx.size()) return SK_ERR_DIM_MISMATCH;
    return plate_force(st, st->bridge_idx.data(), drives, st->bridge_wgt.data(), scale, n, 1);
}

SK_API int amr_inject_plate_force(AcousticAMRState* st, int n, const int32_t* idx,
                                  const float* wgt, float total) {
    if (!st) return SK_ERR_NULL_STATE;                               // :1581-1597
    if (!st->plate_on) return SK_ERR_DIM_MISMATCH;
    if (n <= 0 || !idx || !wgt) return SK_ERR_DIM_MISMATCH;
    int ok = 0;                                      // entries before the first bad node
    while (ok < n && idx[ok] >= 0 && idx[ok] < st->n_pnodes) ++ok;
    const int rc = plate_force(st, idx, wgt, nullptr, total, ok, 0);
    if (rc != SK_OK) return rc;
    return ok == n ? SK_OK : SK_ERR_DIM_MISMATCH;
}

static const float* plate_w(const AcousticAMRState* st) { return st->d_pw[(st->prot + 1) % 3]; }

SK_API int amr_sample_plate_displacement_batch(const AcousticAMRState* cst, int n,
                                               const int* idx4, const float* wgt4, float* out) {
    if (!cst || !idx4 || !wgt4 || !out) return SK_ERR_NULL_STATE;    // :1600-1616
    if (!cst->plate_on) return SK_ERR_DIM_MISMATCH;
    AcousticAMRState* st = const_cast<AcousticAMRState*>(cst);
    int ok = 0;                                      // points before the first bad tap
    while (ok < n) {
        bool good = true;
        for (int k = 0; k < 4; ++k)
            good = good && idx4[4 * ok + k] >= 0 && idx4[4 * ok + k] < st->n_pnodes;
        if (!good) break;
        ++ok;
    }
    if (ok > 0) {
        hipError_t e = ensure_tmp(st, (size_t)ok * 9 * sizeof(float));
        int* di = (int*)st->d_tmp;
        float* dw = (float*)(di + 4 * ok);
        float* dout = dw + 4 * ok;
        if (e == hipSuccess)
            e = hipMemcpyAsync(di, idx4, sizeof(int) * 4 * ok, hipMemcpyHostToDevice, st->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dw, wgt4, sizeof(float) * 4 * ok, hipMemcpyHostToDevice, st->stream);
        if (e == hipSuccess) {
            k_amr_plate_sample4<<<blocks(ok), 256, 0, st->stream>>>(plate_w(st), di, dw, ok, dout);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, dout, sizeof(float) * ok, hipMemcpyDeviceToHost, st->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
        if (e != hipSuccess) return AMR_ERR_DEVICE;
    }
    return ok == n || n <= 0 ? SK_OK : SK_ERR_DIM_MISMATCH;
}

SK_API int amr_sample_plate_weighted(const AcousticAMRState* cst, int n, const int* idx,
                                     const float* wgt, float* out) {
    if (!cst || !idx || !wgt || !out) return SK_ERR_NULL_STATE;      // :1618-1632
    if (!cst->plate_on) return SK_ERR_DIM_MISMATCH;
    AcousticAMRState* st = const_cast<AcousticAMRState*>(cst);
    for (int i = 0; i < n; ++i)
        if (idx[i] < 0 || idx[i] >= st->n_pnodes) return SK_ERR_DIM_MISMATCH;
    const int m = n > 0 ? n : 0;
    hipError_t e = ensure_tmp(st, (size_t)(2 * m + 1) * sizeof(float));
    int* di = (int*)st->d_tmp;
    float* dw = (float*)(di + m);
    float* dout = dw + m;
    if (e == hipSuccess && m)
        e = hipMemcpyAsync(di, idx, sizeof(int) * m, hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess && m)
        e = hipMemcpyAsync(dw, wgt, sizeof(float) * m, hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess) {
        k_amr_plate_weighted<<<1, 1, 0, st->stream>>>(plate_w(st), di, dw, m, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(float), hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    return e == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

SK_API int amr_get_plate_displacement(const AcousticAMRState* cst, float* out, int n) {
    if (!cst || !out) return SK_ERR_NULL_STATE;                      // :1634-1641
    if (!cst->plate_on) return SK_ERR_DIM_MISMATCH;
    if (n != cst->n_pnodes) return SK_ERR_DIM_MISMATCH;
    AcousticAMRState* st = const_cast<AcousticAMRState*>(cst);
    hipError_t e = hipMemcpyAsync(out, plate_w(st), sizeof(float) * n, hipMemcpyDeviceToHost,
                                  st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    return e == hipSuccess ? SK_OK : AMR_ERR_DEVICE;
}

// ── Mic samplers (acoustic_amr.cpp:1647-1792) ───────────────────────────

static hipError_t upload_mics(AcousticAMRState* st) {
    const int M = st->n_mics;
    std::vector<int> I(2 * (M + 1), 0);
    for (int m = 0; m < M; ++m) {
        I[m + 1] = I[m] + (int)st->mp_cell[m].size();
        I[M + 1 + m + 1] = I[M + 1 + m] + (int)st->mv_face[m].size();
    }
    const int np = I[M], nv = I[2 * M + 1];
    std::vector<float> F;
    for (int m = 0; m < M; ++m) {
        I.insert(I.end(), st->mp_cell[m].begin(), st->mp_cell[m].end());
        F.insert(F.end(), st->mp_w[m].begin(), st->mp_w[m].end());
    }
    for (int m = 0; m < M; ++m) I.insert(I.end(), st->mv_face[m].begin(), st->mv_face[m].end());
    for (auto* lst : {&st->mv_wx, &st->mv_wy, &st->mv_wz})
        for (int m = 0; m < M; ++m) F.insert(F.end(), (*lst)[m].begin(), (*lst)[m].end());
    st->o_mvs = M + 1;
    st->o_mpc = 2 * (M + 1);
    st->o_mvf = st->o_mpc + np;
    st->o_mwx = np;
    st->o_mwy = np + nv;
    st->o_mwz = np + 2 * nv;
    drop_graph(st);
    if (st->d_mint) (void)hipFree(st->d_mint);
    if (st->d_mflt) (void)hipFree(st->d_mflt);
    if (st->d_mout) (void)hipFree(st->d_mout);
    st->d_mint = nullptr;
    st->d_mflt = nullptr;
    st->d_mout = nullptr;
    hipError_t e = hipMalloc((void**)&st->d_mint, sizeof(int) * I.size());
    if (e == hipSuccess) e = hipMalloc((void**)&st->d_mflt, sizeof(float) * (F.size() + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&st->d_mout, sizeof(float) * 4 * M);
    if (e == hipSuccess)
        e = hipMemcpyAsync(st->d_mint, I.data(), sizeof(int) * I.size(), hipMemcpyHostToDevice,
                           st->stream);
    if (e == hipSuccess && !F.empty())
        e = hipMemcpyAsync(st->d_mflt, F.data(), sizeof(float) * F.size(), hipMemcpyHostToDevice,
                           st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    ++st->gen;
    return e;
}

SK_API int amr_setup_mic_samplers(AcousticAMRState* st, int n_mics, const float* pos,
                                  const float* axis, int need_v) {
    if (!st || !pos || !axis) return SK_ERR_NULL_STATE;
    if (n_mics <= 0) return SK_ERR_DIM_MISMATCH;
    st->n_mics = n_mics;
    for (auto* l : {&st->mp_cell, &st->mv_face}) l->resize(n_mics);
    for (auto* l : {&st->mp_w, &st->mv_wx, &st->mv_wy, &st->mv_wz}) l->resize(n_mics);
    double mean_vol = 0.0;
    int n_ac = 0;
    float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int i = 0; i < st->nc; ++i) {
        if (!(st->kind[i] & 2)) continue;
        mean_vol += st->volumes[i];
        ++n_ac;
        for (int d = 0; d < 3; ++d) {
            const float c = (float)st->centers[3 * (size_t)i + d];
            lo[d] = c < lo[d] ? c : lo[d];
            hi[d] = c > hi[d] ? c : hi[d];
        }
    }
    if (n_ac > 0) mean_vol /= (double)n_ac;
    const float R = 2.0f * (float)cbrt(mean_vol);
    int rc = SK_OK;
    for (int m = 0; m < n_mics && rc == SK_OK; ++m) {
        float q[3];
        for (int d = 0; d < 3; ++d) q[d] = std::max(lo[d], std::min(hi[d], pos[3 * m + d]));
        auto& pc = st->mp_cell[m];
        auto& pw = st->mp_w[m];
        float wsum = 0.0f;
        const size_t before = pc.size();
        for (int i = 0; i < st->nc; ++i) {
            if (!(st->kind[i] & 2)) continue;
            const float dx = q[0] - (float)st->centers[3 * (size_t)i];
            const float dy = q[1] - (float)st->centers[3 * (size_t)i + 1];
            const float dz = q[2] - (float)st->centers[3 * (size_t)i + 2];
            const float r = sqrtf(dx * dx + dy * dy + dz * dz);
            if (r < R) {
                const float w = (r < 1e-9f) ? 1e9f : 1.0f / r;
                pc.push_back(i);
                pw.push_back(w);
                wsum += w;
            }
        }
        if (pc.empty()) { rc = SK_ERR_DIM_MISMATCH; break; }
        (void)before;
        const float inv = 1.0f / wsum;
        for (float& w : pw) w *= inv;
        if (!need_v) continue;
        float tot = 0.0f;
        for (int f = 0; f < st->nf; ++f) {
            const double* A = &st->centers[3 * (size_t)st->neg[f]];
            const double* B = &st->centers[3 * (size_t)st->pos[f]];
            const float fx = 0.5f * (float)(A[0] + B[0]);
            const float fy = 0.5f * (float)(A[1] + B[1]);
            const float fz = 0.5f * (float)(A[2] + B[2]);
            const float ex = fx - q[0], ey = fy - q[1], ez = fz - q[2];
            const float r2 = sqrtf(ex * ex + ey * ey + ez * ez);
            if (r2 >= R || st->face_open[f] < 0.5) continue;
            float nx = (float)(B[0] - A[0]), ny = (float)(B[1] - A[1]), nz = (float)(B[2] - A[2]);
            const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
            if (nl < 1e-12f) continue;
            nx /= nl;
            ny /= nl;
            nz /= nl;
            const float aw = (float)(st->face_area[f] * st->face_open[f]);
            st->mv_face[m].push_back(f);
            st->mv_wx[m].push_back(nx * aw);
            st->mv_wy[m].push_back(ny * aw);
            st->mv_wz[m].push_back(nz * aw);
            tot += aw;
        }
        if (tot > 1e-20f) {
            const float inv_a = 1.0f / tot;
            for (auto* l : {&st->mv_wx, &st->mv_wy, &st->mv_wz})
                for (float& w : (*l)[m]) w *= inv_a;
        }
    }
    if (upload_mics(st) != hipSuccess) return AMR_ERR_DEVICE;
    return rc;
}

SK_API int amr_sample_mics(const AcousticAMRState* cst, int n_mics, float* op, float* ovx,
                           float* ovy, float* ovz) {
    if (!cst || !op) return SK_ERR_NULL_STATE;                       // :1762-1792
    if (n_mics != cst->n_mics) return SK_ERR_DIM_MISMATCH;
    if (n_mics == 0) return SK_OK;
    AcousticAMRState* st = const_cast<AcousticAMRState*>(cst);
    const int* I = st->d_mint;
    const float* F = st->d_mflt;
    k_amr_mics<<<n_mics, 256, 0, st->stream>>>(
        I, I + st->o_mpc, F, I + st->o_mvs, I + st->o_mvf, F + st->o_mwx, F + st->o_mwy,
        F + st->o_mwz, st->d_p, st->d_v, n_mics, st->d_mout);
    const bool vel = ovx && ovy && ovz;
    std::vector<float> h(4 * (size_t)n_mics);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), st->d_mout, sizeof(float) * h.size(), hipMemcpyDeviceToHost,
                           st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    if (e != hipSuccess) return AMR_ERR_DEVICE;
    for (int m = 0; m < n_mics; ++m) {
        op[m] = h[m];
        if (vel) {
            ovx[m] = h[n_mics + m];
            ovy[m] = h[2 * (size_t)n_mics + m];
            ovz[m] = h[3 * (size_t)n_mics + m];
        }
    }
    return SK_OK;
}

// Extension (no reference counterpart): n_samples audio samples of
// `substeps` steps each with the mic output of every sample recorded on the
// device, one host synchronisation for the whole block.
SK_API int amr_record_mic_samples(AcousticAMRState* st, int n_samples, int substeps,
                                  const float* polar_a, const float* polar_b,
                                  const float* axis_xyz, const float* gain, float rho_c,
                                  float* out, int* n_done) {
    if (!st || !polar_a || !polar_b || !axis_xyz || !gain || !out || !n_done)
        return SK_ERR_NULL_STATE;
    *n_done = 0;
    if (n_samples < 0 || substeps <= 0 || st->n_mics <= 0) return SK_ERR_DIM_MISMATCH;
    if (n_samples > INT_MAX / substeps) return SK_ERR_DIM_MISMATCH;   // step count overflows
    if (n_samples == 0) return SK_OK;
    const int M = st->n_mics;
    std::vector<float> mix(6 * (size_t)M);
    for (int m = 0; m < M; ++m) {
        mix[6 * m] = polar_a[m];
        mix[6 * m + 1] = polar_b[m];
        for (int d = 0; d < 3; ++d) mix[6 * m + 2 + d] = axis_xyz[3 * m + d];
        mix[6 * m + 5] = gain[m];
    }
    const size_t out_n = (size_t)n_samples * M;
    hipError_t e = ensure_tmp(st, sizeof(float) * (mix.size() + out_n));
    float* dmix = (float*)st->d_tmp;
    float* dout = dmix + mix.size();
    if (e == hipSuccess)
        e = hipMemcpyAsync(dmix, mix.data(), sizeof(float) * mix.size(), hipMemcpyHostToDevice,
                           st->stream);
    if (e == hipSuccess) e = hipMemsetAsync(dout, 0, sizeof(float) * out_n, st->stream);
    if (e != hipSuccess) return AMR_ERR_DEVICE;
    const int c0 = st->step_count;
    MicMix mm{substeps, dmix, rho_c, dout};
    const int rc = step_impl(st, n_samples * substeps, nullptr, nullptr, nullptr, &mm);
    if (rc == AMR_ERR_DEVICE) return rc;
    // samples whose steps all completed before any failure
    *n_done = rc == SK_OK ? n_samples : (st->step_count - c0) / substeps;
    e = hipMemcpyAsync(out, dout, sizeof(float) * out_n, hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    if (e != hipSuccess) return AMR_ERR_DEVICE;
    return rc;
}

}  // extern "C"
