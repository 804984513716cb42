// afdtd_batch.hip — many independent rooms of uneven size advanced together.
//
// All rooms live in one allocation per field (room r at element offset
// off[r], each in its own padded z-fastest layout).  A tile table lists every
// (room, y-tile, z-tile, x-segment) work item of every room; ONE launch of the
// batched fused sweep per step covers all of them, followed by ONE launch
// that injects each room's source and gathers its receivers.  For order 2
// the batch advances two steps per launch with the two-step sweep
// (afdtd_fused2.hip), whose in-kernel source / receiver-tap lists for the
// first step are built per room here; the second step's source and
// receivers go through the post-step launch.  fdtd_batch_step captures the
// whole n-step sequence into a hipGraph and launches it once, so the host
// issues one graph launch per call instead of 2 x rooms x steps kernel
// launches.  Rooms are independent: no room reads another's cells.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "afdtd_fused_tile.cuh"

using namespace fused;

struct FdtdBatch {
    int n, order, storage, esz, boundary;
    long long total;
    StencilParams common;
    hipStream_t stream;
    void* P[2];
    void* V[3];
    void* V2[3];
    int cur;
    std::vector<RoomDesc> rooms;
    RoomDesc* d_rooms;
    int* d_tile_room;
    int ntiles, tile_y, seg;
    int schedule;           // tile-table order: 0 room order, 1 largest rooms first
    float* d_alpha;
    float* d_pdamp;
    int* d_status;
    // sources (one per room) and receivers (per_room per room)
    int* d_src;
    float* d_sig;
    int sig_len;
    int per_room;
    int* d_ridx;
    float* d_rwgt;
    float* d_trace;
    int trace_cap, trace_n;
    int step_count;
    // two-step sweep
    int family;             // FDTD_KERNEL_AUTO, _FUSED (single steps) or _FUSED2
    std::vector<int> h_src; // room-local source cells (-1: none)
    std::vector<int> h_ridx;// tap cells of the receivers (global element indices)
    int* d_ls_start;        // per-tile source / tap lists
    int2* d_ls;
    int* d_lt_start;
    int2* d_lt;
    float* d_tapbuf;
    size_t tapcap;
    int lists_dirty, lists_ok;
};

namespace {

struct BatchArgs {
    StencilParams common;
    const RoomDesc* rooms;
    const int* tile_room;
    FusedPtrs f;            // base pointers (room 0)
    const float* alpha;
    const float* pdamp;
    const int* status;
    int seg;
};

template <typename T>
__device__ __forceinline__ FusedPtrs offset_ptrs(const FusedPtrs& b, long long off) {
    FusedPtrs f;
    f.Pin = (const T*)b.Pin + off;
    f.Pout = (T*)b.Pout + off;
    for (int c = 0; c < 3; ++c) {
        f.Vin[c] = (const T*)b.Vin[c] + off;
        f.Vout[c] = (T*)b.Vout[c] + off;
    }
    f.bu = nullptr;
    f.kap = nullptr;
    f.xi = nullptr;
    return f;
}

template <int ORDER, typename T, int TY, int NW>
__global__ void __launch_bounds__(64 * NW)
k_fused_batch(BatchArgs a) {
    if (*a.status) return;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int r = a.tile_room[t];
    const RoomDesc& d = a.rooms[r];
    int local = t - d.tile_base;
    const int ty = local % d.nty; local /= d.nty;
    const int tz = local % d.ntz;
    const int sg = local / d.ntz;
    StencilParams p = a.common;
    p.Nx = d.Nx; p.Ny = d.Ny; p.Nz = d.Nz;
    p.dt = d.dt; p.vel_coeff = d.vel_coeff; p.bulk_coeff = d.bulk_coeff;
    for (int s = 0; s < 6; ++s) { p.gside[s] = d.gside[s]; p.xterm[s] = d.xterm[s]; }
    fused_tile<ORDER, T, false, TY, NW>(p, offset_ptrs<T>(a.f, d.off), a.alpha, a.pdamp, ty, tz,
                                        sg, a.seg);
}

// Per room: add its source sample, then gather its receivers (same order as
// the single-room post-step kernel).
template <typename T>
__global__ void __launch_bounds__(64)
k_batch_post(T* __restrict__ P, const RoomDesc* rooms, const int* __restrict__ src,
             const float* __restrict__ sig, int sig_len, int per_room,
             const int* __restrict__ ridx, const float* __restrict__ rwgt,
             float* __restrict__ trace, int trace_row, int n_rooms, int step, const int* status) {
    typedef Store<T> S;
    if (*status) return;
    const int r = blockIdx.x;
    const long long off = rooms[r].off;
    if (threadIdx.x == 0 && src && step < sig_len && src[r] >= 0) {
        const long long c = off + src[r];
        S::st(P, (int)c, S::ld(P, (int)c) + sig[(size_t)r * sig_len + step]);
    }
    __threadfence_block();
    __syncthreads();
    if (!trace || trace_row < 0) return;
    for (int q = threadIdx.x; q < per_room; q += blockDim.x) {
        const int e = r * per_room + q;
        const int* id = ridx + 8 * e;
        const float* w = rwgt + 8 * e;
        float v = w[0] * S::ld(P, id[0]) + w[1] * S::ld(P, id[1]) + w[2] * S::ld(P, id[2])
                + w[3] * S::ld(P, id[3]) + w[4] * S::ld(P, id[4]) + w[5] * S::ld(P, id[5])
                + w[6] * S::ld(P, id[6]) + w[7] * S::ld(P, id[7]);
        trace[(size_t)trace_row * n_rooms * per_room + e] = v;
    }
}

template <int ORDER, typename T, int TY>
hipError_t launch_batch_t(FdtdBatch* b) {
    BatchArgs a;
    a.common = b->common;
    a.rooms = b->d_rooms;
    a.tile_room = b->d_tile_room;
    a.f.Pin = b->P[b->cur];
    a.f.Pout = b->P[b->cur ^ 1];
    for (int c = 0; c < 3; ++c) { a.f.Vin[c] = b->V[c]; a.f.Vout[c] = b->V2[c]; }
    a.f.bu = nullptr;
    a.f.kap = nullptr;
    a.f.xi = nullptr;
    a.alpha = b->d_alpha;
    a.pdamp = b->d_pdamp;
    a.status = b->d_status;
    a.seg = b->seg;
    constexpr int NW = TY / 2 > 4 ? TY / 2 : 4;     // two rows per thread
    hipLaunchKernelGGL((k_fused_batch<ORDER, T, TY, NW>), dim3(b->ntiles), dim3(64, NW), 0,
                       b->stream, a);
    return hipGetLastError();
}

template <int ORDER, typename T>
hipError_t launch_batch_ty(FdtdBatch* b) {
    switch (b->tile_y) {
        case 8: return launch_batch_t<ORDER, T, 8>(b);
        case 32: return launch_batch_t<ORDER, T, 32>(b);
        default: return launch_batch_t<ORDER, T, 16>(b);
    }
}

template <int ORDER>
hipError_t launch_batch_s(FdtdBatch* b) {
    switch (b->storage) {
        case FDTD_STORE_BF16: return launch_batch_ty<ORDER, bf16_t>(b);
        case FDTD_STORE_F16: return launch_batch_ty<ORDER, f16_t>(b);
        default: return launch_batch_ty<ORDER, float>(b);
    }
}

hipError_t launch_batch_post(FdtdBatch* b, int step, int row) {
    if (!b->d_src && !(b->d_trace && row >= 0)) return hipSuccess;
    void* P = b->P[b->cur];
#define POST(T)                                                                               \
    hipLaunchKernelGGL(k_batch_post<T>, dim3(b->n), dim3(64), 0, b->stream, (T*)P, b->d_rooms, \
                       b->d_src, b->d_sig, b->sig_len, b->per_room, b->d_ridx, b->d_rwgt,     \
                       b->d_trace, row, b->n, step, b->d_status)
    switch (b->storage) {
        case FDTD_STORE_BF16: POST(bf16_t); break;
        case FDTD_STORE_F16: POST(f16_t); break;
        default: POST(float);
    }
#undef POST
    return hipGetLastError();
}

void dfree(void* p) { if (p) (void)hipFree(p); }

int enqueue_batch_step(FdtdBatch* b) {
    hipError_t e = b->order == 4 ? launch_batch_s<4>(b) : launch_batch_s<2>(b);
    if (e != hipSuccess) { afdtd_set_error(hipGetErrorString(e)); return FDTD_ERR_DEVICE; }
    b->cur ^= 1;
    for (int c = 0; c < 3; ++c) std::swap(b->V[c], b->V2[c]);
    int row = -1;
    if (b->d_trace && b->trace_n < b->trace_cap) row = b->trace_n++;
    e = launch_batch_post(b, b->step_count, row);
    if (e != hipSuccess) { afdtd_set_error(hipGetErrorString(e)); return FDTD_ERR_DEVICE; }
    b->step_count++;
    return FDTD_OK;
}

// Whether the batch advances two steps per launch: order 2, a two-step tile
// height (8 or 16 rows), not forced to single steps, lists within the caps.
bool pairs_active(const FdtdBatch* b) {
    return b->order == 2 && (b->tile_y == 8 || b->tile_y == 16) &&
           b->family != FDTD_KERNEL_FUSED && b->lists_ok;
}

// Per-tile source / tap lists of every room for the two-step sweep (list
// index list_base + ty + nty*tz), rebuilt when the tiles, the sources or the
// receivers change.
int build_pair_lists(FdtdBatch* b) {
    if (!b->lists_dirty) return FDTD_OK;
    b->lists_ok = 1;
    if (b->order != 2 || (b->tile_y != 8 && b->tile_y != 16)) {
        b->lists_dirty = 0;
        return FDTD_OK;
    }
    const int TYO = f2_tile_rows(b->tile_y);
    auto by_key = [](const int2& x, const int2& y) { return x.x < y.x || (x.x == y.x && x.y < y.y); };
    std::vector<int> ss, ts;
    std::vector<int2> sflat, tflat;
    const int per = b->per_room;
    std::vector<int> cells, slots;
    for (int r = 0; r < b->n; ++r) {
        const RoomDesc& d = b->rooms[r];
        const int sc = (int)b->h_src.size() == b->n ? b->h_src[r] : -1;
        cells.clear();
        slots.clear();
        for (int q = 0; q < 8 * per && !b->h_ridx.empty(); ++q) {
            const size_t slot = (size_t)8 * r * per + q;
            cells.push_back((int)(b->h_ridx[slot] - d.off));
            slots.push_back((int)slot);
        }
        std::vector<std::vector<int2>> src, tap;
        f2_tile_lists(d.Nx, d.Ny, d.Nz, TYO, 1, &sc, &r, (int)cells.size(), cells.data(),
                      slots.data(), src, tap);
        for (size_t t = 0; t < src.size(); ++t) {
            std::stable_sort(src[t].begin(), src[t].end(), by_key);
            std::stable_sort(tap[t].begin(), tap[t].end(), by_key);
            if ((int)src[t].size() > F2_MAX_SRC || (int)tap[t].size() > F2_MAX_TAP) b->lists_ok = 0;
            ss.push_back((int)sflat.size());
            sflat.insert(sflat.end(), src[t].begin(), src[t].end());
            ts.push_back((int)tflat.size());
            tflat.insert(tflat.end(), tap[t].begin(), tap[t].end());
        }
    }
    ss.push_back((int)sflat.size());
    ts.push_back((int)tflat.size());
    if (sflat.empty()) sflat.push_back(make_int2(-1, 0));
    if (tflat.empty()) tflat.push_back(make_int2(-1, 0));
    auto up = [&](void** dptr, const void* h, size_t bytes) -> hipError_t {
        if (*dptr) { hipError_t e = hipFree(*dptr); *dptr = nullptr; if (e != hipSuccess) return e; }
        hipError_t e = hipMalloc(dptr, bytes);
        if (e != hipSuccess) return e;
        return hipMemcpyAsync(*dptr, h, bytes, hipMemcpyHostToDevice, b->stream);
    };
    AFDTD_CHECK(up((void**)&b->d_ls_start, ss.data(), ss.size() * sizeof(int)));
    AFDTD_CHECK(up((void**)&b->d_ls, sflat.data(), sflat.size() * sizeof(int2)));
    AFDTD_CHECK(up((void**)&b->d_lt_start, ts.data(), ts.size() * sizeof(int)));
    AFDTD_CHECK(up((void**)&b->d_lt, tflat.data(), tflat.size() * sizeof(int2)));
    const size_t ntap = per > 0 ? (size_t)8 * b->n * per : 8;
    if (!b->d_tapbuf || b->tapcap < ntap) {
        dfree(b->d_tapbuf);
        b->d_tapbuf = nullptr;
        AFDTD_CHECK(hipMalloc(&b->d_tapbuf, ntap * sizeof(float)));
        AFDTD_CHECK(hipMemsetAsync(b->d_tapbuf, 0, ntap * sizeof(float), b->stream));
        b->tapcap = ntap;
    }
    AFDTD_CHECK(hipStreamSynchronize(b->stream));
    b->lists_dirty = 0;
    return FDTD_OK;
}

// Steps s and s+1 of every room in one two-step launch: the source and taps
// of step s inside the sweep, those of step s+1 in the post-step launch.
int enqueue_batch_pair(FdtdBatch* b) {
    F2BatchLaunch L;
    L.common = b->common;
    L.Pin = b->P[b->cur];
    L.Pout = b->P[b->cur ^ 1];
    for (int c = 0; c < 3; ++c) {
        L.Vin[c] = b->V[c];
        L.Vout[c] = b->V2[c];
    }
    L.rooms = b->d_rooms;
    L.tile_room = b->d_tile_room;
    L.ntiles = b->ntiles;
    L.seg = b->seg;
    L.tile_y = b->tile_y;
    L.storage = b->storage;
    L.alpha = b->d_alpha;
    L.pdamp = b->d_pdamp;
    L.status = b->d_status;
    L.src_start = b->d_ls_start;
    L.src = b->d_ls;
    L.tap_start = b->d_lt_start;
    L.tap = b->d_lt;
    L.psig = b->d_sig;
    L.plen = b->d_src ? b->sig_len : 0;
    L.step = b->step_count;
    L.tapbuf = b->d_tapbuf;
    L.stream = b->stream;
    hipError_t e = launch_fused2_batch(L);
    if (e != hipSuccess) { afdtd_set_error(hipGetErrorString(e)); return FDTD_ERR_DEVICE; }
    b->cur ^= 1;
    for (int c = 0; c < 3; ++c) std::swap(b->V[c], b->V2[c]);
    int row = -1;
    if (b->d_trace && b->trace_n < b->trace_cap) row = b->trace_n++;
    e = launch_tap_trace(b->d_tapbuf, b->d_rwgt, b->n * b->per_room, b->d_trace, row,
                         b->d_status, b->stream);
    if (e != hipSuccess) { afdtd_set_error(hipGetErrorString(e)); return FDTD_ERR_DEVICE; }
    row = -1;
    if (b->d_trace && b->trace_n < b->trace_cap) row = b->trace_n++;
    e = launch_batch_post(b, b->step_count + 1, row);
    if (e != hipSuccess) { afdtd_set_error(hipGetErrorString(e)); return FDTD_ERR_DEVICE; }
    b->step_count += 2;
    return FDTD_OK;
}

// n steps: pairs while two or more remain and the two-step sweep applies.
int enqueue_batch_steps(FdtdBatch* b, int n_steps) {
    const bool pairs = pairs_active(b);
    for (int s = 0; s < n_steps;) {
        int rc;
        if (pairs && s + 1 < n_steps) {
            rc = enqueue_batch_pair(b);
            s += 2;
        } else {
            rc = enqueue_batch_step(b);
            s += 1;
        }
        if (rc) return rc;
    }
    return FDTD_OK;
}

// (Re)build the tile table for tile height tile_y, x-segment seg and the
// dispatch order `schedule` (0: rooms in order; 1: largest rooms first, so
// the long rooms start first and the tail of the launch holds small ones).
// Each room's tiles stay contiguous (the kernel finds its tile by
// t - tile_base).
int build_tiles(FdtdBatch* b, int tile_y, int seg, int schedule) {
    b->tile_y = tile_y;
    b->seg = seg;
    b->schedule = schedule;
    std::vector<int> order(b->n);
    for (int r = 0; r < b->n; ++r) order[r] = r;
    if (schedule == 1)
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            const RoomDesc &a = b->rooms[x], &c = b->rooms[y];
            return (long long)a.Nx * a.Ny * a.Nz > (long long)c.Nx * c.Ny * c.Nz;
        });
    std::vector<int> tile_room;
    for (int r : order) {
        RoomDesc& d = b->rooms[r];
        d.nty = (d.Ny + tile_y - 1) / tile_y;
        d.ntz = (d.Nz + 63) / 64;
        d.nseg = (d.Nx + seg - 1) / seg;
        d.tile_base = (int)tile_room.size();
        for (int q = 0; q < d.nty * d.ntz * d.nseg; ++q) tile_room.push_back(r);
    }
    int lists = 0;
    for (RoomDesc& d : b->rooms) {
        d.list_base = lists;
        lists += d.nty * d.ntz;
    }
    b->lists_dirty = 1;
    if (b->d_tile_room && (int)tile_room.size() != b->ntiles) {
        dfree(b->d_tile_room);
        b->d_tile_room = nullptr;
    }
    b->ntiles = (int)tile_room.size();
    if (!b->d_tile_room) AFDTD_CHECK(hipMalloc(&b->d_tile_room, sizeof(int) * b->ntiles));
    AFDTD_CHECK(hipMemcpyAsync(b->d_tile_room, tile_room.data(), sizeof(int) * b->ntiles,
                               hipMemcpyHostToDevice, b->stream));
    if (!b->d_rooms) AFDTD_CHECK(hipMalloc(&b->d_rooms, sizeof(RoomDesc) * b->n));
    AFDTD_CHECK(hipMemcpyAsync(b->d_rooms, b->rooms.data(), sizeof(RoomDesc) * b->n,
                               hipMemcpyHostToDevice, b->stream));
    AFDTD_CHECK(hipStreamSynchronize(b->stream));
    return FDTD_OK;
}

}  // namespace

extern "C" {

SK_API void fdtd_batch_destroy(FdtdBatch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (int i = 0; i < 2; ++i) dfree(b->P[i]);
    for (int c = 0; c < 3; ++c) { dfree(b->V[c]); dfree(b->V2[c]); }
    dfree(b->d_rooms); dfree(b->d_tile_room); dfree(b->d_alpha); dfree(b->d_pdamp);
    dfree(b->d_status); dfree(b->d_src); dfree(b->d_sig); dfree(b->d_ridx); dfree(b->d_rwgt);
    dfree(b->d_trace);
    dfree(b->d_ls_start); dfree(b->d_ls); dfree(b->d_lt_start); dfree(b->d_lt);
    dfree(b->d_tapbuf);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

SK_API FdtdBatch* fdtd_batch_create(int n_rooms, const int* dims, float dx, float c,
                                    float rho_air, const float* courant, const float* refl,
                                    int stencil_order, int storage, int tile_y, int seg_x) {
    if (n_rooms <= 0 || !dims || !courant || dx <= 0.0f || c <= 0.0f) {
        afdtd_set_error("fdtd_batch_create: invalid arguments");
        return nullptr;
    }
    if (stencil_order != 2 && stencil_order != 4) return nullptr;
    if (refl && stencil_order != 2) {
        afdtd_set_error("impedance walls require stencil order 2");
        return nullptr;
    }
    FdtdBatch* b = new FdtdBatch();
    b->n = n_rooms;
    b->family = FDTD_KERNEL_AUTO;
    b->lists_ok = 1;
    b->lists_dirty = 1;
    b->order = stencil_order;
    b->storage = storage;
    b->esz = storage == FDTD_STORE_F32 ? 4 : 2;
    b->tile_y = tile_y > 0 ? tile_y : 16;
    b->seg = seg_x > 0 ? seg_x : 32;
    if (b->tile_y != 8 && b->tile_y != 16 && b->tile_y != 32) {
        afdtd_set_error("fdtd_batch_create: tile_y must be 8, 16 or 32");
        delete b;
        return nullptr;
    }
    b->boundary = refl ? FDTD_BOUNDARY_IMPEDANCE : FDTD_BOUNDARY_RIGID_PML;
    StencilParams& sp = b->common;
    memset(&sp, 0, sizeof(sp));
    sp.order = stencil_order;
    sp.n_pml = 0;
    sp.impedance = refl != nullptr;
    sp.dx = dx;
    sp.C1 = 9.0f / 8.0f;
    sp.C2 = -1.0f / 24.0f;
    sp.rho_air = rho_air;
    sp.K0 = rho_air * c * c;
    sp.dtdx = 0.0f;
    sp.sV = 1.0f;
    if (storage == FDTD_STORE_F16) sp.sV = ldexpf(1.0f, (int)lrint(log2((double)rho_air * c)));
    sp.inv_sV = 1.0f / sp.sV;
    const float inv_dx = 1.0f / dx;
    long long off = 0;
    for (int r = 0; r < n_rooms; ++r) {
        RoomDesc d;
        memset(&d, 0, sizeof(d));
        d.Nx = dims[3 * r]; d.Ny = dims[3 * r + 1]; d.Nz = dims[3 * r + 2];
        if (d.Nx <= 0 || d.Ny <= 0 || d.Nz <= 0) { fdtd_batch_destroy(b); return nullptr; }
        d.off = off;
        // Per-room dt from its Courant number; coefficients as the single-room
        // set-up (acoustic_fdtd.cpp:1089, :1176).
        float dt = (float)((double)courant[r] * (double)dx / (double)c);
        d.dt = dt;
        d.vel_coeff = dt / (rho_air * dx);
        d.bulk_coeff = rho_air * c * c * dt * inv_dx;
        const float z0 = sqrtf(rho_air * sp.K0), hk = 0.5f * d.bulk_coeff;
        for (int s = 0; s < 6; ++s) {
            float R = refl ? refl[6 * r + s] : 1.0f;
            d.gside[s] = (1.0f - R) / (1.0f + R);
            d.xterm[s] = hk * (d.gside[s] / z0);
        }
        off += (long long)d.Nx * d.Ny * d.Nz;
        b->rooms.push_back(d);
    }
    b->total = off;
    if (off >= (1LL << 31)) {
        afdtd_set_error("batch too large for 32-bit element indices");
        fdtd_batch_destroy(b);
        return nullptr;
    }
#define BT(expr)                                                                      \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            afdtd_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));       \
            fdtd_batch_destroy(b);                                                    \
            return nullptr;                                                           \
        }                                                                             \
    } while (0)
    BT(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    const size_t bytes = (size_t)off * b->esz;
    for (int i = 0; i < 2; ++i) {
        BT(hipMalloc(&b->P[i], bytes));
        BT(hipMemsetAsync(b->P[i], 0, bytes, b->stream));
    }
    for (int q = 0; q < 3; ++q) {
        BT(hipMalloc(&b->V[q], bytes));
        BT(hipMalloc(&b->V2[q], bytes));
        BT(hipMemsetAsync(b->V[q], 0, bytes, b->stream));
        BT(hipMemsetAsync(b->V2[q], 0, bytes, b->stream));
    }
    if (build_tiles(b, b->tile_y, b->seg, 0) != FDTD_OK) {
        fdtd_batch_destroy(b);
        return nullptr;
    }
    float one = 1.0f, zero = 0.0f;
    BT(hipMalloc(&b->d_alpha, sizeof(float)));
    BT(hipMalloc(&b->d_pdamp, sizeof(float)));
    BT(hipMemcpyAsync(b->d_alpha, &zero, sizeof(float), hipMemcpyHostToDevice, b->stream));
    BT(hipMemcpyAsync(b->d_pdamp, &one, sizeof(float), hipMemcpyHostToDevice, b->stream));
    BT(hipMalloc(&b->d_status, 2 * sizeof(int)));
    BT(hipMemsetAsync(b->d_status, 0, 2 * sizeof(int), b->stream));
    BT(hipStreamSynchronize(b->stream));
#undef BT
    return b;
}

SK_API int fdtd_batch_cells(const FdtdBatch* b, long long* total, int* n_tiles) {
    if (!b) return FDTD_ERR_NULL;
    if (total) *total = b->total;
    if (n_tiles) *n_tiles = b->ntiles;
    return FDTD_OK;
}

// Autotuner: time every (sweep, tile height, x-segment, dispatch order)
// candidate with `trial_steps` graph-captured steps of the batch itself
// (zero fields, sources and recording off, so the timing does not depend on
// the data; an even number for the two-step sweep), keep the fastest and
// leave the batch at step 0 with zero fields.  The sweeps tried are the
// single-step one (tile heights 8, 16, 32) and, for order 2, the two-step
// one (8, 16), unless fdtd_batch_set_kernel fixed one.  Must be called
// before the first step.
SK_API int fdtd_batch_autotune(FdtdBatch* b, int trial_steps, int* best_tile_y, int* best_seg,
                               int* best_schedule, float* best_ms_per_step) {
    if (!b) return FDTD_ERR_NULL;
    if (b->step_count != 0 || trial_steps <= 0) {
        afdtd_set_error("fdtd_batch_autotune: call before the first step, trial_steps > 0");
        return FDTD_ERR_ARG;
    }
    int* src = b->d_src;
    float* trace = b->d_trace;
    b->d_src = nullptr;
    b->d_trace = nullptr;
    const int segs[4] = {16, 32, 64, 128};
    struct Cand { int family, tile_y; };
    std::vector<Cand> cands;
    const int fixed = b->family;
    if (fixed != FDTD_KERNEL_FUSED2)
        for (int ty : {8, 16, 32}) cands.push_back({FDTD_KERNEL_FUSED, ty});
    if (b->order == 2 && fixed != FDTD_KERNEL_FUSED)
        for (int ty : {8, 16}) cands.push_back({FDTD_KERNEL_FUSED2, ty});
    float best = 1e30f;
    int bt = b->tile_y, bs = b->seg, bo = b->schedule, bf = fixed;
    int rc = FDTD_OK;
    for (size_t ic = 0; ic < cands.size() && rc == FDTD_OK; ++ic)
        for (int is = 0; is < 4 && rc == FDTD_OK; ++is)
            for (int o = 0; o < 2 && rc == FDTD_OK; ++o) {
                b->family = cands[ic].family;
                rc = build_tiles(b, cands[ic].tile_y, segs[is], o);
                if (rc) break;
                const int n = cands[ic].family == FDTD_KERNEL_FUSED2 ? trial_steps + (trial_steps & 1)
                                                                      : trial_steps;
                rc = fdtd_batch_step(b, 2, 0, nullptr);               // warm-up
                float ms = 0.0f;
                if (!rc) rc = fdtd_batch_step(b, n, 1, &ms);
                if (!rc && ms / n < best) {
                    best = ms / n;
                    bt = cands[ic].tile_y; bs = segs[is]; bo = o; bf = cands[ic].family;
                }
            }
    b->d_src = src;
    b->d_trace = trace;
    if (rc) { b->family = fixed; return rc; }
    b->family = fixed == FDTD_KERNEL_AUTO ? (bf == FDTD_KERNEL_FUSED2 ? FDTD_KERNEL_AUTO : FDTD_KERNEL_FUSED)
                                          : fixed;
    rc = build_tiles(b, bt, bs, bo);
    if (rc) return rc;
    const size_t bytes = (size_t)b->total * b->esz;
    for (int i = 0; i < 2; ++i) AFDTD_CHECK(hipMemsetAsync(b->P[i], 0, bytes, b->stream));
    for (int q = 0; q < 3; ++q) {
        AFDTD_CHECK(hipMemsetAsync(b->V[q], 0, bytes, b->stream));
        AFDTD_CHECK(hipMemsetAsync(b->V2[q], 0, bytes, b->stream));
    }
    AFDTD_CHECK(hipMemsetAsync(b->d_status, 0, 2 * sizeof(int), b->stream));
    AFDTD_CHECK(hipStreamSynchronize(b->stream));
    b->cur = 0;
    b->step_count = 0;
    b->trace_n = 0;
    if (best_tile_y) *best_tile_y = bt;
    if (best_seg) *best_seg = bs;
    if (best_schedule) *best_schedule = bo;
    if (best_ms_per_step) *best_ms_per_step = best;
    return FDTD_OK;
}

// One source per room: cell_idx[r] (room-local flat index, -1 = none) and
// signals[r*n_samples + n] added after the pressure update of step n.
SK_API int fdtd_batch_set_sources(FdtdBatch* b, const int* cell_idx, const float* signals,
                                  int n_samples) {
    if (!b || !cell_idx || !signals || n_samples <= 0) return FDTD_ERR_NULL;
    for (int r = 0; r < b->n; ++r) {
        const RoomDesc& d = b->rooms[r];
        if (cell_idx[r] >= d.Nx * d.Ny * d.Nz) return FDTD_ERR_DIM;
    }
    dfree(b->d_src); dfree(b->d_sig);
    b->d_src = nullptr; b->d_sig = nullptr;
    AFDTD_CHECK(hipMalloc(&b->d_src, sizeof(int) * b->n));
    AFDTD_CHECK(hipMalloc(&b->d_sig, sizeof(float) * (size_t)b->n * n_samples));
    AFDTD_CHECK(hipMemcpy(b->d_src, cell_idx, sizeof(int) * b->n, hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(b->d_sig, signals, sizeof(float) * (size_t)b->n * n_samples,
                          hipMemcpyHostToDevice));
    b->sig_len = n_samples;
    b->h_src.assign(cell_idx, cell_idx + b->n);
    b->lists_dirty = 1;
    return FDTD_OK;
}

// per_room receivers per room at room-local grid coordinates (n*per_room*3),
// 8-tap trilinear stencils as fdtd_precompute_pressure_samplers.
SK_API int fdtd_batch_set_receivers(FdtdBatch* b, int per_room, const float* xyz) {
    if (!b || !xyz || per_room <= 0) return FDTD_ERR_NULL;
    std::vector<int> idx((size_t)8 * b->n * per_room);
    std::vector<float> wgt((size_t)8 * b->n * per_room);
    for (int r = 0; r < b->n; ++r) {
        const RoomDesc& d = b->rooms[r];
        for (int q = 0; q < per_room; ++q) {
            const size_t e = (size_t)r * per_room + q;
            const float* p = xyz + 3 * e;
            float gx = p[0], gy = p[1], gz = p[2];
            gx = std::min(std::max(gx, 0.0f), (float)(d.Nx - 1));
            gy = std::min(std::max(gy, 0.0f), (float)(d.Ny - 1));
            gz = std::min(std::max(gz, 0.0f), (float)(d.Nz - 1));
            int i0 = (int)gx, i1 = (i0 < d.Nx - 1) ? i0 + 1 : i0;
            int j0 = (int)gy, j1 = (j0 < d.Ny - 1) ? j0 + 1 : j0;
            int k0 = (int)gz, k1 = (k0 < d.Nz - 1) ? k0 + 1 : k0;
            float fx = gx - (float)i0, fy = gy - (float)j0, fz = gz - (float)k0;
            auto I = [&](int i, int j, int k) { return (int)(d.off + k + d.Nz * (j + d.Ny * i)); };
            int* id = &idx[8 * e];
            float* w = &wgt[8 * e];
            id[0] = I(i0, j0, k0); w[0] = (1 - fx) * (1 - fy) * (1 - fz);
            id[1] = I(i1, j0, k0); w[1] = (fx) * (1 - fy) * (1 - fz);
            id[2] = I(i0, j1, k0); w[2] = (1 - fx) * (fy) * (1 - fz);
            id[3] = I(i1, j1, k0); w[3] = (fx) * (fy) * (1 - fz);
            id[4] = I(i0, j0, k1); w[4] = (1 - fx) * (1 - fy) * (fz);
            id[5] = I(i1, j0, k1); w[5] = (fx) * (1 - fy) * (fz);
            id[6] = I(i0, j1, k1); w[6] = (1 - fx) * (fy) * (fz);
            id[7] = I(i1, j1, k1); w[7] = (fx) * (fy) * (fz);
        }
    }
    dfree(b->d_ridx); dfree(b->d_rwgt); dfree(b->d_trace);
    b->d_ridx = nullptr; b->d_rwgt = nullptr; b->d_trace = nullptr;
    b->trace_cap = b->trace_n = 0;
    AFDTD_CHECK(hipMalloc(&b->d_ridx, sizeof(int) * idx.size()));
    AFDTD_CHECK(hipMalloc(&b->d_rwgt, sizeof(float) * wgt.size()));
    AFDTD_CHECK(hipMemcpy(b->d_ridx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
    AFDTD_CHECK(hipMemcpy(b->d_rwgt, wgt.data(), sizeof(float) * wgt.size(), hipMemcpyHostToDevice));
    b->per_room = per_room;
    b->h_ridx = idx;
    b->lists_dirty = 1;
    return FDTD_OK;
}

SK_API int fdtd_batch_record(FdtdBatch* b, int max_steps) {
    if (!b) return FDTD_ERR_NULL;
    if (b->per_room <= 0 || max_steps <= 0) return FDTD_ERR_DIM;
    dfree(b->d_trace);
    b->d_trace = nullptr;
    AFDTD_CHECK(hipMalloc(&b->d_trace, sizeof(float) * (size_t)max_steps * b->n * b->per_room));
    b->trace_cap = max_steps;
    b->trace_n = 0;
    return FDTD_OK;
}

// Advance every room n_steps steps.  use_graph != 0 captures the sequence in
// a hipGraph and launches it once; 0 launches the kernels directly.
// ms_total (optional) receives the device time of the call from HIP events.
SK_API int fdtd_batch_step(FdtdBatch* b, int n_steps, int use_graph, float* ms_total) {
    if (!b) return FDTD_ERR_NULL;
    if (n_steps <= 0) return FDTD_OK;
    int rl = build_pair_lists(b);
    if (rl) return rl;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms_total) {
        AFDTD_CHECK(hipEventCreate(&e0));
        AFDTD_CHECK(hipEventCreate(&e1));
    }
    if (use_graph) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        AFDTD_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_batch_steps(b, n_steps);
        hipError_t ec = hipStreamEndCapture(b->stream, &g);
        if (rc != FDTD_OK) return rc;
        AFDTD_CHECK(ec);
        AFDTD_CHECK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        if (ms_total) AFDTD_CHECK(hipEventRecord(e0, b->stream));
        AFDTD_CHECK(hipGraphLaunch(ge, b->stream));
        if (ms_total) AFDTD_CHECK(hipEventRecord(e1, b->stream));
        AFDTD_CHECK(hipStreamSynchronize(b->stream));
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
    } else {
        if (ms_total) AFDTD_CHECK(hipEventRecord(e0, b->stream));
        int rc = enqueue_batch_steps(b, n_steps);
        if (rc) return rc;
        if (ms_total) AFDTD_CHECK(hipEventRecord(e1, b->stream));
        AFDTD_CHECK(hipStreamSynchronize(b->stream));
    }
    if (ms_total) {
        AFDTD_CHECK(hipEventElapsedTime(ms_total, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return FDTD_OK;
}

SK_API int fdtd_batch_get_traces(const FdtdBatch* b, float* out, int len, int* n_recorded) {
    if (!b || !out) return FDTD_ERR_NULL;
    size_t need = (size_t)b->trace_n * b->n * b->per_room;
    if ((size_t)len < need) return FDTD_ERR_DIM;
    if (need) AFDTD_CHECK(hipMemcpy(out, b->d_trace, sizeof(float) * need, hipMemcpyDeviceToHost));
    if (n_recorded) *n_recorded = b->trace_n;
    return FDTD_OK;
}

SK_API int fdtd_batch_get_pressure(const FdtdBatch* b, int room, float* out, int len) {
    if (!b || !out) return FDTD_ERR_NULL;
    if (room < 0 || room >= b->n) return FDTD_ERR_DIM;
    const RoomDesc& d = b->rooms[room];
    const int N = d.Nx * d.Ny * d.Nz;
    if (len != N) return FDTD_ERR_DIM;
    if (b->storage == FDTD_STORE_F32) {
        AFDTD_CHECK(hipMemcpy(out, (const float*)b->P[b->cur] + d.off, sizeof(float) * N,
                              hipMemcpyDeviceToHost));
        return FDTD_OK;
    }
    float* tmp = nullptr;
    AFDTD_CHECK(hipMalloc(&tmp, sizeof(float) * N));
    const char* base = (const char*)b->P[b->cur] + (size_t)d.off * b->esz;
    AFDTD_CHECK(launch_to_f32(base, b->storage, 1.0f, tmp, N, b->stream, 1));
    AFDTD_CHECK(hipStreamSynchronize(b->stream));
    AFDTD_CHECK(hipMemcpy(out, tmp, sizeof(float) * N, hipMemcpyDeviceToHost));
    dfree(tmp);
    return FDTD_OK;
}

SK_API int fdtd_batch_get_step_count(const FdtdBatch* b) { return b ? b->step_count : 0; }

// FDTD_KERNEL_AUTO (two steps per launch where the two-step sweep applies),
// FDTD_KERNEL_FUSED (single steps) or FDTD_KERNEL_FUSED2 (two-step sweep;
// needs order 2 and tile height 8 or 16).
SK_API int fdtd_batch_set_kernel(FdtdBatch* b, int family) {
    if (!b) return FDTD_ERR_NULL;
    if (family != FDTD_KERNEL_AUTO && family != FDTD_KERNEL_FUSED && family != FDTD_KERNEL_FUSED2) {
        afdtd_set_error("fdtd_batch_set_kernel: unknown kernel family");
        return FDTD_ERR_ARG;
    }
    if (family == FDTD_KERNEL_FUSED2 && (b->order != 2 || (b->tile_y != 8 && b->tile_y != 16))) {
        afdtd_set_error("fdtd_batch_set_kernel: the two-step sweep needs order 2 and tile_y 8 or 16");
        return FDTD_ERR_ARG;
    }
    b->family = family;
    b->lists_dirty = 1;
    return FDTD_OK;
}

// The sweep the next steps use: FDTD_KERNEL_FUSED2 or FDTD_KERNEL_FUSED.
SK_API int fdtd_batch_get_kernel_family(FdtdBatch* b) {
    if (!b) return FDTD_ERR_NULL;
    if (build_pair_lists(b)) return FDTD_ERR_DEVICE;
    return pairs_active(b) ? FDTD_KERNEL_FUSED2 : FDTD_KERNEL_FUSED;
}

}  // extern "C"
