// afdtd_fused.hip — one-pass leapfrog step: 2.5-D tiles marching along x.
//
// A workgroup owns a (TY rows in y) x (64 columns in z) tile and marches an
// x-segment [i0, i1) of it.  Per x-plane q it
//   1. takes the P^n plane q+LA (LA = 3 for order 4, 1 for order 2) from
//      registers into an LDS ring (the load was issued one plane earlier, so
//      HBM latency hides behind one plane of arithmetic);
//   2. updates the Vx face q+FHI of its own columns from the ring (the new
//      Vx faces q-2..q+1 live in a per-thread register queue);
//   3. updates Vy and Vz of plane q on the tile plus the halo faces the
//      divergence needs, into LDS;
//   4. updates P at plane q from the register queue and the LDS faces.
// Each field is read once and written once per step from HBM (32 B per cell
// update in FP32, +8 B for per-cell rho and K), apart from halo re-reads
// that neighbouring tiles share through L2.  Halo faces are recomputed, not
// exchanged, so the velocities are double-buffered (read V^{n-1/2} from one
// buffer, write V^{n+1/2} to the other) and P is double-buffered as in the
// generic path.
//
// The per-face and per-cell arithmetic is the reference's, in the same
// order (acoustic_fdtd.cpp:1107-1113 gradient/velocity, :894-925 divergence,
// :1198 pressure), so results equal the generic kernels and the oracle bit
// for bit in FP32.
//
// With FP32 storage the sweep also carries the reference's cell geometry
// (the GEOM build): walls and plate cells keep their pressure, closed faces
// are set to 0 after their update (acoustic_fdtd.cpp:1146-1149), plate faces
// take the plate velocity (w - w_prev)/dt (:1151-1158) and the divergence
// taps are weighted by the face fractions (:894-925).  A code byte per cell
// says which of these apply, so a cell costs one byte more than in an empty
// room; only faces whose fraction is not 1 read their fraction, and only
// plate faces the plate displacements, so the O(surface) geometry adds
// O(surface) traffic.
//
// Workgroup order is XCD-aware: blocks are dealt round-robin to the 8 XCDs,
// so the block index is remapped to give each XCD a contiguous range of
// tiles (y fastest), which keeps halo rows of neighbouring tiles in the same
// L2.
#include "afdtd_fused_tile.cuh"

using namespace fused;

namespace {

struct FusedArgs {
    StencilParams p;
    FusedPtrs f;
    const float* alpha;   // alpha(d)
    const float* pdamp;   // 1/(1 + alpha(d) dt)
    const int* status;
    GeoPtrs g;            // cell geometry (GEOM build)
    int seg;              // x-segment length
    int nty, ntz, nseg;   // tile counts
    // side job (GEOM build, side_n > 0): the last workgroup forms a weighted
    // plate sum instead of a tile
    const float* side_w;
    const int* side_idx;
    const float* side_wgt;
    int side_n, side_nplate, side_tag;
    float* side_out;
    int run_on;           // the sweep runs while the status word equals this
};

// The side job: products w[idx[i]] * wgt[i] (0 for an index outside the
// plate) formed by the workgroup in chunks, summed by one thread in index
// order from +0; the sum and the tag go to side_out.
__device__ void side_sum(const FusedArgs& a) {
    constexpr int CH = 2048;
    __shared__ float sb[CH];
    const int tid = threadIdx.x + 64 * threadIdx.y, nt = 64 * blockDim.y;
    float acc = 0.0f;
    for (int base = 0; base < a.side_n; base += CH) {
        const int m = min(CH, a.side_n - base);
        for (int i = tid; i < m; i += nt) {
            const int p = a.side_idx[base + i];
            sb[i] = (p < 0 || p >= a.side_nplate) ? 0.0f : a.side_w[p] * a.side_wgt[base + i];
        }
        __syncthreads();
        if (tid == 0) {
            int i = 0;
            for (; i + 8 <= m; i += 8) {
                float v[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) v[q] = sb[i + q];
#pragma unroll
                for (int q = 0; q < 8; ++q) acc += v[q];
            }
            for (; i < m; ++i) acc += sb[i];
        }
        __syncthreads();
    }
    if (tid == 0) {
        a.side_out[0] = acc;
        ((int*)a.side_out)[1] = a.side_tag;
    }
}

// Minimum waves per SIMD the register allocation must allow (occupancy).
#ifndef AFDTD_FUSED_WPE
#define AFDTD_FUSED_WPE 1
#endif
#ifndef AFDTD_FUSED_GEO_WPE
#define AFDTD_FUSED_GEO_WPE 1
#endif

template <int ORDER, typename T, bool HET, int TY, int NW, bool GEOM>
__global__ void __launch_bounds__(64 * NW)
__attribute__((amdgpu_waves_per_eu(GEOM ? AFDTD_FUSED_GEO_WPE : AFDTD_FUSED_WPE)))
k_fused(FusedArgs a) {
    if (*a.status != a.run_on) return;
    int ntile = gridDim.x;
    if constexpr (GEOM) {
        if (a.side_n > 0) {
            if (blockIdx.x == gridDim.x - 1) {
                side_sum(a);
                return;
            }
            ntile -= 1;
        }
    }
    int t = xcd_remap(blockIdx.x, ntile);
    const int ty = t % a.nty; t /= a.nty;
    const int tz = t % a.ntz;
    const int sg = t / a.ntz;
    fused_tile<ORDER, T, HET, TY, NW, GEOM>(a.p, a.f, a.alpha, a.pdamp, ty, tz, sg, a.seg, a.g);
}

template <int ORDER, typename T, bool HET, int TY, bool GEOM>
hipError_t launch_t(AcousticFDTDState* st, const FusedArgs& a0) {
    constexpr int NW = TY / 2 > 4 ? TY / 2 : 4;     // two rows per thread
    FusedArgs a = a0;
    a.nty = (st->Ny + TY - 1) / TY;
    a.ntz = (st->Nz + TZ - 1) / TZ;
    a.nseg = (st->Nx + a.seg - 1) / a.seg;
    const int nb = a.nty * a.ntz * a.nseg + (GEOM && a.side_n > 0 ? 1 : 0);
    hipLaunchKernelGGL((k_fused<ORDER, T, HET, TY, NW, GEOM>), dim3(nb), dim3(64, NW), 0,
                       st->stream, a);
    return hipGetLastError();
}

template <int ORDER, typename T, bool HET, bool GEOM = false>
hipError_t launch_ty(AcousticFDTDState* st, const FusedArgs& a, int ty) {
    switch (ty) {
        case 8: return launch_t<ORDER, T, HET, 8, GEOM>(st, a);
        case 32: return launch_t<ORDER, T, HET, 32, GEOM>(st, a);
        default: return launch_t<ORDER, T, HET, 16, GEOM>(st, a);
    }
}

template <int ORDER, typename T>
hipError_t launch_het(AcousticFDTDState* st, const FusedArgs& a, int ty) {
    return st->hetero ? launch_ty<ORDER, T, true>(st, a, ty) : launch_ty<ORDER, T, false>(st, a, ty);
}

template <int ORDER>
hipError_t launch_store(AcousticFDTDState* st, const FusedArgs& a, int ty) {
    if (st->fused_geo)
        return st->hetero ? launch_ty<ORDER, float, true, true>(st, a, ty)
                          : launch_ty<ORDER, float, false, true>(st, a, ty);
    switch (st->storage) {
        case FDTD_STORE_BF16: return launch_het<ORDER, bf16_t>(st, a, ty);
        case FDTD_STORE_F16: return launch_het<ORDER, f16_t>(st, a, ty);
        case FDTD_STORE_BFP16: return launch_het<ORDER, bfp16_t>(st, a, ty);
        case FDTD_STORE_F16S: return launch_het<ORDER, f16s_t>(st, a, ty);
        default: return launch_het<ORDER, float>(st, a, ty);
    }
}

}  // namespace

bool fused_supported(const AcousticFDTDState* st) {
    return !st->has_solid && !st->has_frac && st->nact == 0 && st->n_pml <= 64;
}

bool fused_geo_supported(const AcousticFDTDState* st) {
    return st->storage == FDTD_STORE_F32 && st->n_pml <= 64 &&
           (size_t)st->N * sizeof(float) < 0x80000000u;
}

hipError_t launch_fused_step(AcousticFDTDState* st) { return launch_fused_step_if(st, 0); }

hipError_t launch_fused_step_if(AcousticFDTDState* st, int run_on) {
    if (!st->V2[0]) return hipErrorInvalidValue;
    FusedArgs a;
    a.run_on = run_on;
    a.p = st->sp;
    a.f.Pin = st->P[st->cur];
    a.f.Pout = st->P[st->cur ^ 1];
    for (int c = 0; c < 3; ++c) {
        a.f.Vin[c] = st->V[c];
        a.f.Vout[c] = st->V2[c];
    }
    a.f.bu = st->d_bu;
    a.f.kap = st->d_kap;
    a.f.xi = st->d_xi;
    a.alpha = st->d_alpha;
    a.pdamp = st->d_pdamp;
    a.status = st->d_status;
    a.g = GeoPtrs{};
    if (st->fused_geo) {
        if (!st->d_geo) return hipErrorInvalidValue;
        a.g.code = st->d_geo;
        a.g.frac[0] = st->d_fx; a.g.frac[1] = st->d_fy; a.g.frac[2] = st->d_fz;
        a.g.vpl = st->d_vpl;
        a.g.ncells = (uint32_t)st->N;
    }
    a.side_n = 0;
    if (st->fused_geo && st->side_n > 0 && st->side_out) {
        a.side_w = st->d_w[0];                 // the displacement this step ends with
        a.side_idx = st->side_idx;
        a.side_wgt = st->side_wgt;
        a.side_n = st->side_n;
        a.side_nplate = st->Nplate;
        a.side_tag = st->side_step;
        a.side_out = st->side_out;
    }
    // 8-row tiles for the geometry build: its larger register set leaves
    // three workgroups of 4 waves per CU, against one of 8 waves at 16 rows
    int ty = st->fused_tile_y > 0 ? st->fused_tile_y : (st->fused_geo ? 8 : 16);
    a.seg = st->fused_seg_x > 0 ? st->fused_seg_x : 32;
    if (st->fused_seg_x <= 0) {
        // small grids: shorter x-segments (down to 4 planes) until the launch
        // has two workgroups per CU; each cell's arithmetic does not depend on
        // the segment (a 718 K-cell body: 32 planes 98 us, 8 planes 36 us,
        // 4 planes 27 us, 2 planes 32 us per step)
        const long tiles = (long)((st->Ny + ty - 1) / ty) * ((st->Nz + TZ - 1) / TZ);
        while (a.seg > 4 && tiles * ((st->Nx + a.seg - 1) / a.seg) < 512) a.seg /= 2;
    }
    if (a.seg > st->Nx) a.seg = st->Nx;
    hipError_t e = st->order == 4 ? launch_store<4>(st, a, ty) : launch_store<2>(st, a, ty);
    if (e == hipSuccess) {
        for (int c = 0; c < 3; ++c) {
            void* t = st->V[c];
            st->V[c] = st->V2[c];
            st->V2[c] = t;
        }
    }
    return e;
}
