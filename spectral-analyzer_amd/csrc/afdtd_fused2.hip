// afdtd_fused2.hip — two leapfrog steps per sweep (order-2 stencil).
//
// The single-step fused sweep (afdtd_fused.hip) reads and writes every field
// once per step, so its speed is the streaming rate of that byte mix.  This
// sweep advances two steps, n -> n+2, per pass over HBM: a workgroup owns a
// 16-row x 64-column tile and marches an x-segment of it, and per x-plane q
//   stage A (step n -> n+1) updates V^{n+1/2} and P^{n+1} on the tile plus a
//           one-cell ring around it (the "extended" tile: rows 0..17,
//           columns -1..64, own cells are rows 1..16, columns 0..63), from
//           P^n planes staged in an LDS ring (rows -1..18, columns -2..65);
//   stage B (step n+1 -> n+2) trails stage A by one plane: at plane q-1 it
//           updates V^{n+3/2} of the faces of the own cells and P^{n+2} of
//           the own cells from P^{n+1} (LDS and registers) and V^{n+1/2}
//           (registers of the thread that computed it in stage A), and
//           writes them to HBM.
// P^{n+1} and V^{n+1/2} never go to HBM, so one pass costs one read and one
// write of each field for two cell-updates.  The ring cells of stage A are
// recomputed by the neighbouring tiles with identical arithmetic.
//
// Work mapping (640 threads = 10 wavefronts of 64 lanes):
//   waves 0..8 ("main"): lane = column 0..63, wave w = extended rows 2w, 2w+1;
//   wave 9 ("halo"):  lanes 0..15 the cells (e, -1), e = 1..16, lanes 16..31
//                      the cells (e, 64), lanes 32/33 the z-face between
//                      columns -1 and 0 on rows 0 and 17, and every lane one
//                      y-face of the face row between rows -1 and 0.
//
// Arithmetic is the single-step sweep's, in the same order
// (acoustic_fdtd.cpp:1107-1113 gradient/velocity, :894-925 divergence,
// :1198 pressure); with 16-bit storage the stage-A results are rounded to the
// storage type exactly where the single-step path stores and reloads them,
// and the point sources of step n+1 are added after that rounding as the
// post-step kernel does.  So results equal two single-step sweeps bit for
// bit.  Receiver taps of step n+1 are written to a small buffer by the
// thread that owns the tap cell; the stability check never falls on step n+1
// (the host schedules pairs so that it does not).
#include "afdtd_fused_tile.cuh"

#include <string.h>

#include <algorithm>

using fused::FusedPtrs;
using fused::xcd_remap;

namespace {

namespace f2 {
constexpr int TZ = 64;           // own columns (one per lane)
constexpr int GC = TZ + 4;       // P^n region columns -2..65
constexpr int NSP = 4;           // P^n ring: planes q-1..q+1 read, q+2 refilled (power of two)
constexpr int NSB = 8;           // 1/rho ring: planes q-2..q+1 read, q+2 refilled
constexpr int EC = TZ + 2;       // P^{n+1} columns -1..64
// Tile geometry for R extended rows per main thread.
template <int R>
struct Geo2 {
    static constexpr int NWM = R == 2 ? 9 : 10;   // main waves
    static constexpr int TYO = R * NWM - 2;       // own rows: 16 (R = 2) or 8 (R = 1)
    static constexpr int NW = NWM + 1;            // + the halo wave
    static constexpr int NT = 64 * NW;
    static constexpr int EY = TYO + 2;            // extended rows
    static constexpr int GR = EY + 2;             // P^n region rows -1..EY
    static constexpr int GPL = GR * GC + 1;       // + one spare word (stores of idle lanes)
    static constexpr int EPL = EY * EC;
    static constexpr int NHL = 4 * GR;            // region halo-column elements per plane
    static constexpr int LR = (GR + NW - 1) / NW; // region rows loaded per wave
    static constexpr int NS = R > 2 ? R : 2;      // old-value slots per field
};
constexpr int MAXS = F2_MAX_SRC; // point-source entries cached per tile
constexpr int MAXT = F2_MAX_TAP; // receiver-tap entries cached per tile
// entry key: plane << 12 | extended row << 7 | (extended column + 1)
__host__ __device__ __forceinline__ int ekey(int plane, int e, int c) {
    return (plane << 12) | (e << 7) | (c + 1);
}
// Buffer access: a wave-uniform base in a scalar descriptor plus a 32-bit
// per-lane byte offset (no 64-bit address arithmetic per lane).  `bytes`
// bounds the offsets: reads beyond it return 0, writes beyond it are
// dropped.  Stores of FP32 fields are non-temporal (written once per sweep).
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             0x00020000);
}
template <typename Raw> __device__ __forceinline__ Raw bld(Rsrc r, uint32_t off);
template <> __device__ __forceinline__ float bld<float>(Rsrc r, uint32_t off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ uint16_t bld<uint16_t>(Rsrc r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
}
template <> __device__ __forceinline__ f16_t bld<f16_t>(Rsrc r, uint32_t off) {
    return __builtin_bit_cast(f16_t, (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
}
__device__ __forceinline__ void bst(Rsrc r, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, 0, 2);
}
__device__ __forceinline__ void bst(Rsrc r, uint32_t off, uint16_t v) {
    __builtin_amdgcn_raw_buffer_store_b16(v, r, off, 0, 0);
}
__device__ __forceinline__ void bst(Rsrc r, uint32_t off, f16_t v) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, off, 0, 0);
}
// Per-lane condition held as a VGPR word (all ones or zero) and applied with
// a bitwise AND.  Kept as 64-bit SGPR lane masks, the many per-lane
// conditions of the sweep exceed the scalar register budget and are spilled
// into VGPR lanes and reloaded on every plane; the empty asm makes the word
// opaque, so the compiler does not turn it back into a lane mask.
__device__ __forceinline__ uint32_t vmask(bool c) {
    uint32_t m = c ? 0xFFFFFFFFu : 0u;
    asm volatile("" : "+v"(m));
    return m;
}
__device__ __forceinline__ float keep(float v, uint32_t m) {
    return __uint_as_float(__float_as_uint(v) & m);
}
// LDS index helpers (row / column in extended-tile coordinates)
__device__ __forceinline__ int gix(int r, int c) { return (r + 1) * GC + (c + 2); }
__device__ __forceinline__ int p1ix(int e, int c) { return e * EC + (c + 1); }
// Value of lane l-1 of the wavefront (lane 0 gets `at0`): the z-face to the
// left of a cell, which its left neighbour computed (DPP wave_shr:1; every
// lane of the wavefront must be active).
__device__ __forceinline__ float lane_left(float v, float at0) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, at0),
                                                                 __builtin_bit_cast(int, v),
                                                                 0x138, 0xF, 0xF, false));
}
}  // namespace f2

// Minimum waves per SIMD the register allocation must allow.  Without
// per-cell media a workgroup needs 48 KB of LDS, and two share a CU (the
// limit measured on MI355X is about 54 KB each) when their waves fit the
// register file too: two 10-wave workgroups can put six waves on one SIMD,
// so the budget is 512 / 6 -> 80 VGPRs.  With per-cell media (larger 1/rho
// ring, 93 KB) one workgroup per CU is the limit anyway.
#ifndef AFDTD_F2_WPE
#define AFDTD_F2_WPE 1
#endif
#ifndef AFDTD_F2_WPE_HOMOG
#define AFDTD_F2_WPE_HOMOG 6
#endif

struct Pair2Args {
    StencilParams p;
    FusedPtrs f;
    const float* alpha;
    const float* pdamp;
    const int* status;
    int seg, nty, ntz, nseg;
    const int* src_start;   // per tile t = ty + nty*tz: entries [src_start[t], src_start[t+1])
    const int2* src;        // (key, source index), sorted by key then source index
    const int* tap_start;
    const int2* tap;        // (key, tap slot), sorted by key
    const float* psig;
    int plen;
    int step;               // global index of the first of the two steps
    float* tapbuf;          // P^{n+1} at the receiver taps
};

// One work item: the tile (ty, tz) of x-segment sg of the grid described by
// p (dimensions, coefficients) and fp (field arrays); `tile` selects its
// source and tap lists.  PML: the grid has an absorbing layer (PML = false
// compiles the layer's damping out; the host picks the variant from n_pml).
template <typename T, bool HET, int R, bool PML>
__device__ __forceinline__ void fused2_tile(const Pair2Args& a, const StencilParams& p,
                                            const FusedPtrs& fp, int ty, int tz, int sg,
                                            int tile) {
    using namespace f2;
    typedef Geo2<R> G;
    constexpr int NWM = G::NWM, TYO = G::TYO, NT = G::NT, EY = G::EY, GPL = G::GPL,
                  EPL = G::EPL, NHL = G::NHL, LR = G::LR, NS = G::NS;
    typedef Store<T> S;
    typedef typename S::Raw RawT;
    __shared__ float sP[NSP * GPL];
    __shared__ float sB[HET ? NSB * GPL : 1];
    // Double-buffered by plane parity: one iteration writes the faces (or
    // P^{n+1}) of one plane while it reads those of the plane before.  A
    // cell's own faces and those of the rows of the same thread stay in
    // registers, the z-face to its left comes from the neighbouring lane, so
    // only the faces crossing a wavefront go through LDS: the y-face above a
    // wavefront's first row (sAyX / sByX, slot w for wave w; slot 0 of sAyX
    // is the face row between rows -1 and 0) and the z-face between columns
    // -1 and 0 (sAzC / sBzC, by row).
    constexpr int AXN = (NWM + 1) * TZ, ACN = EY, BXN = (NWM + 1) * TZ, BCN = TYO + 2;
    __shared__ float sP1[2 * EPL];
    __shared__ float sAyX[2 * AXN];
    __shared__ float sAzC[2 * ACN];
    __shared__ float sByX[2 * BXN];
    __shared__ float sBzC[2 * BCN];
    __shared__ float sAlpha[65], sPd[65];
    __shared__ int sSrcKey[MAXS], sTapKey[MAXT], sTapSlot[MAXT];
    __shared__ float sAmp[MAXS];


    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz, sx = Ny * Nz;
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int tid = lane + 64 * w;
    const int j0 = ty * TYO, k0 = tz * TZ;
    const int i0 = sg * a.seg, i1 = min(i0 + a.seg, Nx);
    const int npml = PML ? p.n_pml : 0;
    if (npml > 0)
        for (int d = tid; d <= npml && d < 65; d += NT) { sAlpha[d] = a.alpha[d]; sPd[d] = a.pdamp[d]; }
    const float vcoef = p.vel_coeff, dtdx = p.dtdx, dt = p.dt, kap0 = p.bulk_coeff;
    const float isv = p.inv_sV, sv = p.sV;
    const bool loadXi = HET && p.impedance;

    const T* __restrict__ Pin = (const T*)fp.Pin;
    T* __restrict__ Pout = (T*)fp.Pout;
    const T* __restrict__ Vxi = (const T*)fp.Vin[0];
    const T* __restrict__ Vyi = (const T*)fp.Vin[1];
    const T* __restrict__ Vzi = (const T*)fp.Vin[2];
    T* __restrict__ Vxo = (T*)fp.Vout[0];
    T* __restrict__ Vyo = (T*)fp.Vout[1];
    T* __restrict__ Vzo = (T*)fp.Vout[2];
    const float* __restrict__ bu = fp.bu;
    const float* __restrict__ kapA = fp.kap;
    const float* __restrict__ xiA = fp.xi;

    // ── helpers ──
    // Loads take a per-plane base (uniform; the plane index is clamped into the
    // grid) and a per-lane byte offset fixed for the whole sweep (0 for a lane
    // whose cell lies outside the grid).  A value loaded for an invalid plane
    // or lane only reaches faces and cells that are masked to zero or never
    // stored, so no per-load select is needed.  Stores use the same offsets
    // (see OOB below for cells outside the grid).
    const int esz = (int)sizeof(T);
    const uint32_t pbT = (uint32_t)sx * esz, pbF = (uint32_t)sx * 4;   // plane bytes
    auto cq = [&](int q) __attribute__((always_inline)) { return min(max(q, 0), Nx - 1); };
    auto rT = [&](const T* F, size_t b) __attribute__((always_inline)) { return rsrc(F + b, pbT); };
    auto rF = [&](const float* F, size_t b) __attribute__((always_inline)) { return rsrc(F + b, pbF); };
    auto ldT = [&](Rsrc r, uint32_t boff) __attribute__((always_inline)) { return bld<RawT>(r, boff); };
    auto ldF = [&](Rsrc r, uint32_t boff) __attribute__((always_inline)) { return bld<float>(r, boff); };
    // store of a velocity (isV: scaled for 16-bit storage) or a pressure
    auto stT = [&](T* F, int plane, uint32_t boff, float v, bool isV) __attribute__((always_inline)) {
        bst(rT(F, (size_t)plane * sx), boff, S::raw_of((S::exact || !isV) ? v : v * sv));
    };
    // stored-and-reloaded value of a velocity (rndv) / pressure (rndp)
    // under 16-bit storage; identity for FP32
    auto vel = [&](RawT x) __attribute__((always_inline)) { const float f = S::cvt(x); return S::exact ? f : f * isv; };
    auto rndv = [&](float x) __attribute__((always_inline)) { return S::exact ? x : S::cvt(S::raw_of(x * sv)) * isv; };
    auto rndp = [&](float x) __attribute__((always_inline)) { return S::exact ? x : S::cvt(S::raw_of(x)); };
    auto pd = [&](int x, int n) __attribute__((always_inline)) { return min(pml_depth(x, n, npml), 64); };
    auto damp = [&](int dL, int dR) __attribute__((always_inline)) { return face_damp(sAlpha, dL, dR, dt); };
    auto depth3 = [&](int i, int j, int k) __attribute__((always_inline)) { return imax3(pd(i, Nx), pd(j, Ny), pd(k, Nz)); };
    auto cellv = [&](int j, int k) __attribute__((always_inline)) { return j >= 0 && j < Ny && k >= 0 && k < Nz; };
    // pressure update of one cell (impedance wall when on the domain side)
    auto pres = [&](float pc, float dv, float kap, float xi, int i, int j, int k, bool edge) __attribute__((always_inline)) {
        float out;
        if (p.impedance && edge) {
            const float x = HET ? xi : wall_xi(p.xterm, i, j, k, Nx, Ny, Nz);
            out = impedance_update(pc, dv, kap, x);
        } else {
            out = pc - kap * dv;
        }
        if (npml > 0) out = out * sPd[depth3(i, j, k)];
        return out;
    };

    // ── region loads (P^n and 1/rho of one plane into the ring) ──
    // main part: wave w loads region rows LR*w-1 .. LR*w+LR-2 (columns
    // 0..63); halo columns -2, -1, 64, 65 of the GR rows: threads tid < NHL.
    uint32_t rgT[LR + 1], rgF[LR + 1];
    int ldLds[LR + 1];
#pragma unroll
    for (int m = 0; m < LR; ++m) {
        const int rr = LR * w - 1 + m;                // waves beyond row EY idle
        const int j = j0 - 1 + rr, k = k0 + lane;
        const int off = (rr <= EY && cellv(j, k)) ? j * Nz + k : 0;
        rgT[m] = (uint32_t)(off * esz);
        rgF[m] = (uint32_t)(off * 4);
        ldLds[m] = rr <= EY ? gix(rr, lane) : GPL - 1;   // idle lanes: the spare word
    }
    {
        const int rr = -1 + (tid >> 2), cs = tid & 3;
        const int cc = cs == 0 ? -2 : cs == 1 ? -1 : cs == 2 ? 64 : 65;
        const int j = j0 - 1 + rr, k = k0 + cc;
        const int off = (tid < NHL && cellv(j, k)) ? j * Nz + k : 0;
        rgT[LR] = (uint32_t)(off * esz);
        rgF[LR] = (uint32_t)(off * 4);
        ldLds[LR] = tid < NHL ? gix(rr, cc) : GPL - 1;
    }
    // P^n and 1/rho of one region plane, staged in registers (unconverted)
    struct Stage {
        RawT p[LR + 1];
        float b[LR + 1];
    };
    auto load_plane = [&](int pl) __attribute__((always_inline)) {
        Stage st;
        const size_t b = (size_t)cq(pl) * sx;
        const Rsrc rp = rT(Pin, b);
        const Rsrc rb = rF(HET ? bu : (const float*)Pin, b);
#pragma unroll
        for (int m = 0; m <= LR; ++m) {
            st.p[m] = ldT(rp, rgT[m]);
            st.b[m] = HET ? ldF(rb, rgF[m]) : 0.0f;
        }
        return st;
    };
    auto store_plane = [&](int pl, const Stage& st) __attribute__((always_inline)) {
#pragma unroll
        for (int m = 0; m <= LR; ++m) {
            sP[(pl & (NSP - 1)) * GPL + ldLds[m]] = S::cvt(st.p[m]);
            if (HET) sB[(pl & (NSB - 1)) * GPL + ldLds[m]] = st.b[m];
        }
    };
    auto Pn = [&](int pl, int r, int c) __attribute__((always_inline)) { return sP[(pl & (NSP - 1)) * GPL + gix(r, c)]; };
    auto Bn = [&](int pl, int r, int c) __attribute__((always_inline)) { return sB[(pl & (NSB - 1)) * GPL + gix(r, c)]; };

    // ── thread roles ──
    const bool mainW = w < NWM;
    const int k = k0 + lane;
    const bool kval = k < Nz;
    // main: extended rows e = R*w + r
    int me[R], mj[R], moff[R];
    bool mcell[R], mvy[R], mvz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        me[r] = R * w + r;
        mj[r] = j0 - 1 + me[r];
        moff[r] = mj[r] * Nz + k;
        mcell[r] = mainW && cellv(mj[r], k);
        mvy[r] = mainW && kval && mj[r] >= 0 && mj[r] <= Ny - 2;
        mvz[r] = mainW && mj[r] >= 0 && mj[r] < Ny && k <= Nz - 2;
    }
    // halo wave roles
    const bool hH = !mainW && lane < 2 * TYO;            // cell (he, hc)
    // lanes >= 2*TYO have no cell; they get (1, -1) so that the unconditional
    // stage-A x-face reads stay inside the LDS tiles (result discarded)
    const int he = lane < TYO ? lane + 1 : (lane < 2 * TYO ? lane - TYO + 1 : 1);
    const int hc = (lane < TYO || lane >= 2 * TYO) ? -1 : 64;
    const int hj = j0 - 1 + he, hk = k0 + hc;
    const int hoff = hj * Nz + hk;
    const bool hcell = hH && cellv(hj, hk);
    // Vy faces at rows hj-1, hj (column hk); Vz faces at columns hk-1, hk (row hj)
    const bool hvy0 = hH && hk >= 0 && hk < Nz && hj - 1 >= 0 && hj - 1 <= Ny - 2;
    const bool hvy1 = hH && hk >= 0 && hk < Nz && hj >= 0 && hj <= Ny - 2;
    const bool hvz0 = hH && hj >= 0 && hj < Ny && hk - 1 >= 0 && hk - 1 <= Nz - 2;
    const bool hvz1 = hH && hj >= 0 && hj < Ny && hk >= 0 && hk <= Nz - 2;
    // T1: z-face between columns -1 and 0 on rows 0 / 17
    const bool hT1 = !mainW && (lane == 2 * TYO || lane == 2 * TYO + 1);
    const int t1e = lane == 2 * TYO ? 0 : EY - 1;
    const int t1j = j0 - 1 + t1e;
    const int t1off = t1j * Nz + k0 - 1;
    const bool t1v = hT1 && t1j >= 0 && t1j < Ny && k0 - 1 >= 0 && k0 - 1 <= Nz - 2;
    // T4: y-face between rows -1 and 0 (global face j0-2) at column lane
    const int t4j = j0 - 2;
    const int t4off = t4j * Nz + k;
    const bool t4v = !mainW && kval && t4j >= 0 && t4j <= Ny - 2;
    // domain-side cells (impedance walls) on a y or z side; x sides per plane
    bool medge[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        medge[r] = mcell[r] && (mj[r] == 0 || mj[r] == Ny - 1 || k == 0 || k == Nz - 1);
    const bool hedge = hcell && (hj == 0 || hj == Ny - 1 || hk == 0 || hk == Nz - 1);
    // With per-cell media (one workgroup per CU, VGPRs to spare) the same
    // conditions are used inside the sweep as VGPR words; the homogeneous
    // build, held to 80 VGPRs, keeps the lane masks.
    uint32_t mcellM[R], mvyM[R], mvzM[R], medgeM[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if constexpr (HET) {
            mcellM[r] = vmask(mcell[r]);
            mvyM[r] = vmask(mvy[r]);
            mvzM[r] = vmask(mvz[r]);
            medgeM[r] = vmask(medge[r]);
        } else {
            mcellM[r] = mvyM[r] = mvzM[r] = medgeM[r] = 0u;
        }
    }
    // halo-wave conditions, likewise
    uint32_t hcellM = 0u, hvyM[2] = {0u, 0u}, hvzM[2] = {0u, 0u}, t1vM = 0u, t4vM = 0u,
             hedgeM = 0u;
    if constexpr (HET) {
        hcellM = vmask(hcell);
        hvyM[0] = vmask(hvy0);
        hvyM[1] = vmask(hvy1);
        hvzM[0] = vmask(hvz0);
        hvzM[1] = vmask(hvz1);
        t1vM = vmask(t1v);
        t4vM = vmask(t4v);
        hedgeM = vmask(hedge);
    }
    auto pick = [&](float v, bool c, uint32_t m) __attribute__((always_inline)) {
        if constexpr (HET) return keep(v, m);
        else return c ? v : 0.0f;
    };
    auto edge_or_x = [&](int r, bool xside) __attribute__((always_inline)) {
        if constexpr (HET) return (medgeM[r] | (mcellM[r] & (xside ? ~0u : 0u))) != 0u;
        else return medge[r] || (xside && mcell[r]);
    };

    // Byte offsets of the old-value loads (one set, meaning by role):
    //   main: oT[r] / oF[r] the cell (mj[r], k), xF[r] its wall sum (edges only)
    //   halo: oT[0] the cell (hj, hk), oT[1] (hj-1, hk), oT[2] (hj, hk-1) or the
    //         T1 face, oT[3] the T4 face; oF[0] / xF[0] the cell
    // xD: added to xF on x-side planes, where every cell is a wall cell.
    // oT of a cell outside the grid is OOB, past every plane's bytes: loads
    // there return 0 and stores are dropped, so stores need no per-lane
    // condition (faces of a valid cell that lie outside the grid are stored
    // as 0 into their unused padding slots).
    constexpr uint32_t OOB = 0x80000000u;
    uint32_t oT[R > 4 ? R : 4], oF[NS], xF[NS], xD[NS];
#pragma unroll
    for (int r = 0; r < NS; ++r) oF[r] = xF[r] = xD[r] = 0u;
    oT[1] = 0u;
    if (mainW) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int c = mcell[r] ? moff[r] : 0;
            oT[r] = mcell[r] ? (uint32_t)(c * esz) : OOB;
            oF[r] = (uint32_t)(c * 4);
            xF[r] = medge[r] ? oF[r] : 0u;
            xD[r] = oF[r] - xF[r];
        }
#pragma unroll
        for (int r = R; r < 4; ++r) oT[r] = 0u;
    } else {
        const int c0 = hcell ? hoff : 0;
        const int c1 = (hH && cellv(hj - 1, hk)) ? hoff - Nz : 0;
        const int c2 = hT1 ? (cellv(t1j, k0 - 1) ? t1off : 0)
                           : ((hH && cellv(hj, hk - 1)) ? hoff - 1 : 0);
        const int c3 = cellv(t4j, k) ? t4off : 0;
        oT[0] = hcell ? (uint32_t)(c0 * esz) : OOB;
        oT[1] = (uint32_t)(c1 * esz);
        oT[2] = (uint32_t)(c2 * esz);
        oT[3] = (uint32_t)(c3 * esz);
        oF[0] = (uint32_t)(c0 * 4);
        xF[0] = hedge ? oF[0] : 0u;
        xD[0] = oF[0] - xF[0];
    }
    // Per-slot offsets of the old-value loads, so that both roles issue the
    // same loads (slot x1 reads Vx for main waves and Vy for the halo wave's
    // T4 face; the descriptor is chosen per wave):
    //   slots x[0..NS) y[0..NS) z[0..NS) (T fields), kappa / xi via lf, xF, xD
    uint32_t lo[3 * NS], lf[NS];
    if (mainW) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            lo[r] = lo[NS + r] = lo[2 * NS + r] = oT[r < 4 ? r : 0];
            lf[r] = oF[r];
        }
    } else {
#pragma unroll
        for (int r = 0; r < 3 * NS; ++r) lo[r] = oT[0];
        lo[1] = oT[3];
        lo[NS] = oT[1];
        lo[2 * NS] = oT[2];
#pragma unroll
        for (int r = 0; r < NS; ++r) lf[r] = oF[0];
    }

    // Old values of plane q: V^{n-1/2} (stage-A faces of q) and the kappa /
    // wall sums of plane q-1 (stage-A pressure of q-1, one iteration later).
    // Main waves: x[r], y[r], z[r], kk[r], xi[r] of row r.  The halo wave
    // uses x[0] (Vx), x[1] (the T4 face), y[0..1] (Vy), z[0..1] (Vz / T1),
    // kk[0], xi[0] (kk[1], xi[1] are loaded for the same cell and unused).
    struct Old {
        RawT x[NS], y[NS], z[NS];
        float kk[NS], xi[NS];
    };
    auto load_old = [&](int q) __attribute__((always_inline)) {
        Old o;
        const size_t b = (size_t)cq(q) * sx;
        const size_t bk = (size_t)cq(q - 1) * sx;
        const bool xe = q - 1 == 0 || q - 1 == Nx - 1;
        const Rsrc X = rT(Vxi, b), Y = rT(Vyi, b), Z = rT(Vzi, b);
        const Rsrc X1 = mainW ? X : Y;
        const Rsrc K = rF(HET ? kapA : (const float*)Vxi, bk);
        // without impedance walls the wall-sum descriptor is empty: its loads
        // return 0 and touch no memory
        const Rsrc XI = rsrc((loadXi ? xiA : (const float*)Vxi) + bk, loadXi ? pbF : 0u);
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            o.x[r] = ldT(r == 1 ? X1 : X, lo[r]);
            o.y[r] = ldT(Y, lo[NS + r]);
            o.z[r] = ldT(Z, lo[2 * NS + r]);
        }
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            // every cell of an x-side plane is a wall cell
            o.kk[r] = (HET && r < R) ? ldF(K, lf[r]) : 0.0f;
            o.xi[r] = (HET && r < R) ? ldF(XI, xF[r] + (xe ? xD[r] : 0u)) : 0.0f;
        }
        return o;
    };

    // ── stage A: the x-face between planes q and q+1 at the thread's cells ──
    // main: cells (me[r], lane); halo: cell (he, hc).
    auto vxA = [&](int q, const Old& o, float* out) __attribute__((always_inline)) {
        const bool fv = q >= 0 && q <= Nx - 2;
        if (mainW) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float coef = HET ? (0.5f * (Bn(q, me[r], lane) + Bn(q + 1, me[r], lane))) * dtdx : vcoef;
                float nv = vel(o.x[r]) - coef * (Pn(q + 1, me[r], lane) - Pn(q, me[r], lane));
                if (npml > 0) nv = nv * damp(depth3(q, mj[r], k), depth3(q + 1, mj[r], k));
                out[r] = fv ? pick(nv, mcell[r], mcellM[r]) : 0.0f;
            }
        } else {
            const float coef = HET ? (0.5f * (Bn(q, he, hc) + Bn(q + 1, he, hc))) * dtdx : vcoef;
            float nv = vel(o.x[0]) - coef * (Pn(q + 1, he, hc) - Pn(q, he, hc));
            if (npml > 0) nv = nv * damp(depth3(q, hj, hk), depth3(q + 1, hj, hk));
            out[0] = fv ? pick(nv, hcell, hcellM) : 0.0f;
#pragma unroll
            for (int r = 1; r < R; ++r) out[r] = 0.0f;
        }
    };

    // ── per-thread pipeline registers (iteration q, after its updates) ──
    // Stage A: Vx faces q (xa0), q-1 (xa1), q-2 (xa2); Vy / Vz of planes q,
    // q-1, q-2 (y0..y2, z0..z2); P^{n+1} of planes q-1, q-2, q-3 (pa0..pa2).
    // Stage B: Vx^{n+3/2} faces q-2 (xb0), q-3 (xb1), q-4 (xb2).  kappa / wall
    // sums of planes q-2, q-3 for stage-B pressure (kb1, kb2, xb1w, xb2w;
    // those of plane q-1 arrive with `o`).
    float xa0[R], xa1[R], xa2[R], y0[R], y1[R], y2[R], z0[R], z1[R], z2[R];
    float pa0[R], pa1[R], pa2[R], xb0[R], xb1[R], xb2[R];
    float yb0[R], yb1[R], zb0[R], zb1[R];   // stage-B y / z faces of planes q-2, q-3
    float kb1[R], kb2[R], xw1[R], xw2[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        xa0[r] = xa1[r] = xa2[r] = y0[r] = y1[r] = y2[r] = z0[r] = z1[r] = z2[r] = 0.0f;
        pa0[r] = pa1[r] = pa2[r] = xb0[r] = xb1[r] = xb2[r] = 0.0f;
        yb0[r] = yb1[r] = zb0[r] = zb1[r] = 0.0f;
        kb1[r] = kb2[r] = xw1[r] = xw2[r] = 0.0f;
    }
    // Halo-wave pipeline: Vy faces hj-1, hj and Vz faces hk-1, hk of plane q
    // (stage-A faces) and of plane q-1 (stage-A pressure), and the z-face
    // between columns -1 and 0 of plane q-2 (stage-B face).  With two rows
    // per main thread they ride in the second-row slots of the main
    // pipeline arrays, which the halo wave does not use, and rotate with
    // them: hy0 -> y0[1], hy1 -> z0[1], hz0 -> xb0[1], hz1 -> pa0[1] (so
    // pa2[1] is hz1 of plane q-2).
    float hq[4] = {0.0f, 0.0f, 0.0f, 0.0f}, hp[4] = {0.0f, 0.0f, 0.0f, 0.0f}, hz2s = 0.0f;
    auto hset = [&](float a0, float a1, float b0, float b1) __attribute__((always_inline)) {
        if constexpr (R >= 2) {
            y0[1] = a0; z0[1] = a1; xb0[1] = b0; pa0[1] = b1;
        } else {
            hq[0] = a0; hq[1] = a1; hq[2] = b0; hq[3] = b1;
        }
    };
    auto hprev = [&](int i) __attribute__((always_inline)) {   // plane q-1 values
        if constexpr (R >= 2) return i == 0 ? y1[1] : i == 1 ? z1[1] : i == 2 ? xb1[1] : pa1[1];
        else return hp[i];
    };
    auto hz2v = [&]() __attribute__((always_inline)) {
        if constexpr (R >= 2) return pa2[1];
        else return hz2s;
    };

    // The tile's source and tap entries (sorted by plane) and the source
    // amplitudes of step n+1, cached in LDS; walked with a uniform cursor.
    const int s0 = a.src_start[tile], tp0 = a.tap_start[tile];
    const int ns = a.step < a.plen ? min(a.src_start[tile + 1] - s0, MAXS) : 0;
    const int nt = min(a.tap_start[tile + 1] - tp0, MAXT);
    for (int m = tid; m < ns; m += NT) {
        const int2 en = a.src[s0 + m];
        sSrcKey[m] = en.x;
        sAmp[m] = a.psig[(size_t)en.y * a.plen + a.step];
    }
    for (int m = tid; m < nt; m += NT) {
        const int2 en = a.tap[tp0 + m];
        sTapKey[m] = en.x;
        sTapSlot[m] = en.y;
    }
    int scur = 0, tcur = 0;            // uniform cursors into the cached lists
    const int qf = i0 - 1;             // first plane of stage A (one-plane ring in x)

    // One iteration, one barrier: stage-A faces of plane q, stage-A pressure
    // of plane q-1, stage-B faces of plane q-2 and stage-B pressure of plane
    // q-3.  Each reads only LDS data written before the previous barrier (the
    // face and P^{n+1} buffers alternate by plane parity, the P^n ring holds
    // planes q-2..q+2) or registers of the same thread written earlier in the
    // iteration.  On entry `o` holds the old values of plane q and `st` the
    // staged region plane q+2.  The iteration first uses them up (stage-A
    // faces, ring refill, kappa / wall sums into the pipeline) and then
    // loads the next plane's into the same registers, so no loaded value is
    // copied while its load may still be in flight; the loads have the rest
    // of the iteration to arrive.
    auto plane = [&](int q, Old& o, Stage& st) __attribute__((always_inline)) {
        const int qa = q - 1, qb = q - 2, qc = q - 3;
        const bool qv = q >= 0 && q < Nx;
        const bool qbv = qb >= 0 && qb < Nx;
        float* sAyW = sAyX + (q & 1) * AXN;
        float* sAzW = sAzC + (q & 1) * ACN;
        const float* sAyR = sAyX + (qa & 1) * AXN;
        const float* sAzR = sAzC + (qa & 1) * ACN;

        // ════ stage-A faces of plane q ════
        if (q <= i1) {
            vxA(q, o, xa0);
            if (mainW) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int e = me[r];
                    {   // y-face between rows e and e+1, column lane
                        const float coef = HET ? (0.5f * (Bn(q, e, lane) + Bn(q, e + 1, lane))) * dtdx : vcoef;
                        float nv = vel(o.y[r]) - coef * (Pn(q, e + 1, lane) - Pn(q, e, lane));
                        if (npml > 0) nv = nv * damp(depth3(q, mj[r], k), depth3(q, mj[r] + 1, k));
                        y0[r] = qv ? pick(nv, mvy[r], mvyM[r]) : 0.0f;
                        if (r == R - 1) sAyW[(w + 1) * TZ + lane] = y0[r];
                    }
                    {   // z-face between columns lane and lane+1, row e
                        const float coef = HET ? (0.5f * (Bn(q, e, lane) + Bn(q, e, lane + 1))) * dtdx : vcoef;
                        float nv = vel(o.z[r]) - coef * (Pn(q, e, lane + 1) - Pn(q, e, lane));
                        if (npml > 0) nv = nv * damp(depth3(q, mj[r], k), depth3(q, mj[r], k + 1));
                        z0[r] = qv ? pick(nv, mvz[r], mvzM[r]) : 0.0f;
                    }
                }
            } else {
                if (hH) {
                    float hy[2], hz[2];
#pragma unroll
                    for (int u = 0; u < 2; ++u) {   // y-faces between rows he-1+u and he+u
                        const int e = he - 1 + u;
                        const float coef = HET ? (0.5f * (Bn(q, e, hc) + Bn(q, e + 1, hc))) * dtdx : vcoef;
                        float nv = vel(o.y[u]) - coef * (Pn(q, e + 1, hc) - Pn(q, e, hc));
                        if (npml > 0) nv = nv * damp(depth3(q, hj - 1 + u, hk), depth3(q, hj + u, hk));
                        hy[u] = qv ? pick(nv, u ? hvy1 : hvy0, hvyM[u]) : 0.0f;
                    }
#pragma unroll
                    for (int u = 0; u < 2; ++u) {   // z-faces between columns hc-1+u and hc+u
                        const int c = hc - 1 + u;
                        const float coef = HET ? (0.5f * (Bn(q, he, c) + Bn(q, he, c + 1))) * dtdx : vcoef;
                        float nv = vel(o.z[u]) - coef * (Pn(q, he, c + 1) - Pn(q, he, c));
                        if (npml > 0) nv = nv * damp(depth3(q, hj, hk - 1 + u), depth3(q, hj, hk + u));
                        hz[u] = qv ? pick(nv, u ? hvz1 : hvz0, hvzM[u]) : 0.0f;
                    }
                    if (hc < 0) sAzW[he] = hz[1];
                    hset(hy[0], hy[1], hz[0], hz[1]);
                }
                if (hT1) {
                    const float coef = HET ? (0.5f * (Bn(q, t1e, -1) + Bn(q, t1e, 0))) * dtdx : vcoef;
                    float nv = vel(o.z[0]) - coef * (Pn(q, t1e, 0) - Pn(q, t1e, -1));
                    if (npml > 0) nv = nv * damp(depth3(q, t1j, k0 - 1), depth3(q, t1j, k0));
                    sAzW[t1e] = qv ? pick(nv, t1v, t1vM) : 0.0f;
                }
                {   // T4: y-face between rows -1 and 0
                    const float coef = HET ? (0.5f * (Bn(q, -1, lane) + Bn(q, 0, lane))) * dtdx : vcoef;
                    float nv = vel(o.x[1]) - coef * (Pn(q, 0, lane) - Pn(q, -1, lane));
                    if (npml > 0) nv = nv * damp(depth3(q, t4j, k), depth3(q, t4j + 1, k));
                    sAyW[lane] = qv ? pick(nv, t4v, t4vM) : 0.0f;
                }
            }
            // ring refill: plane q+2 (its P slot last held plane q-2, read
            // up to the previous iteration; its 1/rho slot plane q-6)
            store_plane(q + 2, st);
        }
        float kA[NS], xA[NS];   // kappa / wall sums of plane q-1
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            kA[r] = o.kk[r];
            xA[r] = o.xi[r];
        }
        o = load_old(q + 1);
        st = load_plane(q + 3);

        // ════ stage-A pressure of plane qa = q-1 ════
        if (qa >= qf && qa <= i1) {
            const bool xe = qa == 0 || qa == Nx - 1;
            float* p1s = sP1 + (qa & 1) * EPL;
            if (mainW) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int e = me[r];
                    const float yd = r > 0 ? y1[r > 0 ? r - 1 : 0] : sAyR[w * TZ + lane];
                    const float dv = (xa1[r] - xa2[r]) + (y1[r] - yd)
                                   + (z1[r] - lane_left(z1[r], sAzR[e]));
                    const float kap = HET ? kA[r] : kap0;
                    const float xi = loadXi ? xA[r] : 0.0f;
                    const float out = pres(Pn(qa, e, lane), dv, kap, xi, qa, mj[r], k,
                                           edge_or_x(r, xe));
                    pa0[r] = rndp(out);
                }
            } else if (hH) {
                const float dv = (xa1[0] - xa2[0]) + (hprev(1) - hprev(0)) + (hprev(3) - hprev(2));
                const float kap = HET ? kA[0] : kap0;
                const float xi = loadXi ? xA[0] : 0.0f;
                const float out = pres(Pn(qa, he, hc), dv, kap, xi, qa, hj, hk,
                                       HET ? (hedgeM | (hcellM & (xe ? ~0u : 0u))) != 0u
                                           : hedge || (xe && hcell));
                pa0[0] = rndp(out);
            }
            // point sources of step n+1 (after the rounding, in source order)
            while (scur < ns && (__builtin_amdgcn_readfirstlane(sSrcKey[scur]) >> 12) < qa) ++scur;
            for (; scur < ns; ++scur) {
                const int key = __builtin_amdgcn_readfirstlane(sSrcKey[scur]);
                if ((key >> 12) != qa) break;
                const int e = (key >> 7) & 31, c = (key & 127) - 1;
                const float amp = sAmp[scur];
                if (mainW) {
                    if (e / R == w && c == lane) {
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            if (e % R == r) pa0[r] = rndp(pa0[r] + amp);
                    }
                } else if (hH && e == he && c == hc) {
                    pa0[0] = rndp(pa0[0] + amp);
                }
            }
            if (mainW) {
#pragma unroll
                for (int r = 0; r < R; ++r) p1s[p1ix(me[r], lane)] = pa0[r];
            } else if (hH) {
                p1s[p1ix(he, hc)] = pa0[0];
            }
            // receiver taps of step n+1 (own planes and cells only)
            while (tcur < nt && (__builtin_amdgcn_readfirstlane(sTapKey[tcur]) >> 12) < qa) ++tcur;
            for (; tcur < nt; ++tcur) {
                const int key = __builtin_amdgcn_readfirstlane(sTapKey[tcur]);
                if ((key >> 12) != qa) break;
                const int e = (key >> 7) & 31, c = (key & 127) - 1;
                if (mainW && qa >= i0 && qa < i1 && e / R == w && c == lane) {
                    float v = pa0[0];
#pragma unroll
                    for (int r = 1; r < R; ++r)
                        if (e % R == r) v = pa0[r];
                    a.tapbuf[sTapSlot[tcur]] = v;
                }
            }
        }

        // ════ stage-B faces of plane qb = q-2 ════
        if (qb >= i0 - 1 && qb <= i1 - 1) {
            const bool fvb = qb >= 0 && qb <= Nx - 2;
            const bool ownq = qb >= i0;
            const float* p1b = sP1 + (qb & 1) * EPL;
            float* sByW = sByX + (qb & 1) * BXN;
            float* sBzW = sBzC + (qb & 1) * BCN;
            if (mainW) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int e = me[r];
                    const bool ownr = e >= 1 && e <= TYO;
                    if (ownr) {   // x-face between planes qb and qb+1
                        const float coef = HET ? (0.5f * (Bn(qb, e, lane) + Bn(qb + 1, e, lane))) * dtdx : vcoef;
                        float nv = rndv(xa2[r]) - coef * (pa0[r] - pa1[r]);
                        if (npml > 0) nv = nv * damp(depth3(qb, mj[r], k), depth3(qb + 1, mj[r], k));
                        nv = fvb ? pick(nv, mcell[r], mcellM[r]) : 0.0f;
                        xb0[r] = nv;
                        if (ownq && fvb) stT(Vxo, qb, oT[r], nv, true);
                    }
                    if (ownq) {
                        if (e <= TYO) {   // y-face between rows e and e+1
                            const float coef = HET ? (0.5f * (Bn(qb, e, lane) + Bn(qb, e + 1, lane))) * dtdx : vcoef;
                            float nv = rndv(y2[r]) - coef * (p1b[p1ix(e + 1, lane)] - p1b[p1ix(e, lane)]);
                            if (npml > 0) nv = nv * damp(depth3(qb, mj[r], k), depth3(qb, mj[r] + 1, k));
                            nv = qbv ? pick(nv, mvy[r], mvyM[r]) : 0.0f;
                            yb0[r] = nv;
                            if (r == R - 1) sByW[(w + 1) * TZ + lane] = nv;
                            if (ownr && qbv) stT(Vyo, qb, oT[r], nv, true);
                        }
                        if (ownr) {   // z-face between columns lane and lane+1
                            const float coef = HET ? (0.5f * (Bn(qb, e, lane) + Bn(qb, e, lane + 1))) * dtdx : vcoef;
                            float nv = rndv(z2[r]) - coef * (p1b[p1ix(e, lane + 1)] - p1b[p1ix(e, lane)]);
                            if (npml > 0) nv = nv * damp(depth3(qb, mj[r], k), depth3(qb, mj[r], k + 1));
                            nv = qbv ? pick(nv, mvz[r], mvzM[r]) : 0.0f;
                            zb0[r] = nv;
                            if (qbv) stT(Vzo, qb, oT[r], nv, true);
                        }
                    }
                }
            } else if (hH && hc < 0 && ownq) {   // z-face between columns -1 and 0, row he
                const float coef = HET ? (0.5f * (Bn(qb, he, -1) + Bn(qb, he, 0))) * dtdx : vcoef;
                float nv = rndv(hz2v()) - coef * (p1b[p1ix(he, 0)] - pa1[0]);
                if (npml > 0) nv = nv * damp(depth3(qb, hj, k0 - 1), depth3(qb, hj, k0));
                sBzW[he] = qbv ? pick(nv, hvz1, hvzM[1]) : 0.0f;
            }
        }

        // ════ stage-B pressure of plane qc = q-3 ════
        if (qc >= i0 && qc <= i1 - 1 && mainW) {
            const bool xec = qc == 0 || qc == Nx - 1;
            const float* sByR = sByX + (qc & 1) * BXN;
            const float* sBzR = sBzC + (qc & 1) * BCN;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int e = me[r];
                if (e < 1 || e > TYO) continue;
                const float yd = r > 0 ? yb1[r > 0 ? r - 1 : 0] : sByR[w * TZ + lane];
                const float dv = (xb1[r] - xb2[r]) + (yb1[r] - yd)
                               + (zb1[r] - lane_left(zb1[r], sBzR[e]));
                const float out = pres(pa2[r], dv, HET ? kb2[r] : kap0, xw2[r], qc, mj[r], k,
                                       edge_or_x(r, xec));
                stT(Pout, qc, oT[r], out, false);
            }
        }
        __syncthreads();

        // ── rotate the pipeline ──
#pragma unroll
        for (int r = 0; r < R; ++r) {
            xa2[r] = xa1[r];
            xa1[r] = xa0[r];
            y2[r] = y1[r];
            y1[r] = y0[r];
            z2[r] = z1[r];
            z1[r] = z0[r];
            pa2[r] = pa1[r];
            pa1[r] = pa0[r];
            xb2[r] = xb1[r];
            xb1[r] = xb0[r];
            yb1[r] = yb0[r];
            zb1[r] = zb0[r];
            kb2[r] = kb1[r];
            kb1[r] = kA[r];
            xw2[r] = xw1[r];
            xw1[r] = xA[r];
        }
        if constexpr (R == 1) {
            hz2s = hp[3];
#pragma unroll
            for (int i = 0; i < 4; ++i) hp[i] = hq[i];
        }
    };

    // ── prologue: ring planes i0-2 .. i0, staged plane i0+1, Vx face i0-2 ──
    for (int pl = qf - 1; pl <= qf + 1; ++pl) store_plane(pl, load_plane(pl));
    Stage sa = load_plane(qf + 2);
    {
        const Old o0 = load_old(qf - 1);
        __syncthreads();
        vxA(qf - 1, o0, xa1);
    }
    Old oa = load_old(qf);
    // iterations q = i0-1 .. i1+2 (stage-B pressure covers planes i0 .. i1-1)
#pragma unroll 1
    for (int q = qf; q <= i1 + 2; ++q) plane(q, oa, sa);
}

template <typename T, bool HET, int R, bool PML>
__global__ void __launch_bounds__(f2::Geo2<R>::NT)
__attribute__((amdgpu_waves_per_eu(HET ? AFDTD_F2_WPE : AFDTD_F2_WPE_HOMOG)))
k_fused2(Pair2Args a) {
    if (*a.status) return;
    // Each XCD takes a contiguous range of work items (xcd_remap); within a
    // segment the items run in bands of up to four z-tiles, z fastest, that
    // sweep y, so the workgroups resident on one XCD at a time form a compact
    // (y, z) block whose shared halo lines are read once into its L2.
    int t = xcd_remap(blockIdx.x, gridDim.x);
    const int per_sg = a.nty * a.ntz;
    const int sg = t / per_sg;
    t -= sg * per_sg;
    const int band = t / (a.nty * 4);                   // bands of four z-tiles
    const int gz = min(4, a.ntz - 4 * band);            // z-tiles in this band
    t -= band * a.nty * 4;
    const int ty = t / gz;
    const int tz = 4 * band + t % gz;
    const int tile = ty + a.nty * tz;
    fused2_tile<T, HET, R, PML>(a, a.p, a.f, ty, tz, sg, tile);
}

// Batched rooms: work item t belongs to room tile_room[t]; the room's
// dimensions, coefficients and field offset replace the shared ones.  Rooms
// are homogeneous (no per-cell media); a room's work items run in y-fastest
// order.
template <typename T, int R, bool PML>
__global__ void __launch_bounds__(f2::Geo2<R>::NT)
__attribute__((amdgpu_waves_per_eu(AFDTD_F2_WPE_HOMOG)))
k_fused2_batch(Pair2Args a, const RoomDesc* __restrict__ rooms, const int* __restrict__ tile_room) {
    if (*a.status) return;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const RoomDesc& d = rooms[tile_room[t]];
    int local = t - d.tile_base;
    const int ty = local % d.nty;
    local /= d.nty;
    const int tz = local % d.ntz;
    const int sg = local / d.ntz;
    StencilParams p = a.p;
    p.Nx = d.Nx;
    p.Ny = d.Ny;
    p.Nz = d.Nz;
    p.dt = d.dt;
    p.vel_coeff = d.vel_coeff;
    p.bulk_coeff = d.bulk_coeff;
    for (int s = 0; s < 6; ++s) {
        p.gside[s] = d.gside[s];
        p.xterm[s] = d.xterm[s];
    }
    FusedPtrs f;
    f.Pin = (const T*)a.f.Pin + d.off;
    f.Pout = (T*)a.f.Pout + d.off;
    for (int c = 0; c < 3; ++c) {
        f.Vin[c] = (const T*)a.f.Vin[c] + d.off;
        f.Vout[c] = (T*)a.f.Vout[c] + d.off;
    }
    f.bu = nullptr;
    f.kap = nullptr;
    f.xi = nullptr;
    fused2_tile<T, false, R, PML>(a, p, f, ty, tz, sg, d.list_base + ty + d.nty * tz);
}

// Receiver traces of the intermediate step from the tap buffer, in the
// post-step kernel's order (w0*p0 + w1*p1 + ... + w7*p7).
__global__ void __launch_bounds__(256)
k_tap_trace(const float* __restrict__ tapbuf, const float* __restrict__ rwgt, int nrec,
            float* __restrict__ trace, int trace_row, const int* status) {
    if (status[0]) return;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += gridDim.x * blockDim.x) {
        const float* t = tapbuf + 8 * r;
        const float* w = rwgt + 8 * r;
        const float v = w[0] * t[0] + w[1] * t[1] + w[2] * t[2] + w[3] * t[3] + w[4] * t[4]
                      + w[5] * t[5] + w[6] * t[6] + w[7] * t[7];
        trace[(size_t)trace_row * nrec + r] = v;
    }
}

template <typename T, bool HET, bool PML>
hipError_t launch2p(AcousticFDTDState* st, const Pair2Args& a, int nb) {
    if (st->f2_rows == 1)
        hipLaunchKernelGGL((k_fused2<T, HET, 1, PML>), dim3(nb), dim3(64, f2::Geo2<1>::NW), 0,
                           st->stream, a);
    else
        hipLaunchKernelGGL((k_fused2<T, HET, 2, PML>), dim3(nb), dim3(64, f2::Geo2<2>::NW), 0,
                           st->stream, a);
    return hipGetLastError();
}
template <typename T, bool HET>
hipError_t launch2(AcousticFDTDState* st, const Pair2Args& a, int nb) {
    return a.p.n_pml > 0 ? launch2p<T, HET, true>(st, a, nb) : launch2p<T, HET, false>(st, a, nb);
}

}  // namespace

bool fused2_supported(const AcousticFDTDState* st) {
    return (st->order == 2 && fused_supported(st)) || fused4_supported(st);
}

// Per-tile source and tap list entries of one grid (host side).  Tile
// t = ty + nty*tz covers extended rows j0-1..j0+TYO and columns k0-1..k0+64
// of every plane: a source goes to every tile whose stage A computes its cell
// (main rows 0..TYO+1 x columns 0..63, halo cells rows 1..TYO x columns -1
// and 64), a tap to the tile that owns its cell.
void f2_tile_lists(int Nx, int Ny, int Nz, int TYO, int nsrc, const int* src_cell,
                   const int* src_id, int ntap, const int* tap_cell, const int* tap_slot,
                   std::vector<std::vector<int2>>& src, std::vector<std::vector<int2>>& tap) {
    using namespace f2;
    (void)Nx;
    const int EY = TYO + 2;
    const int nty = (Ny + TYO - 1) / TYO, ntz = (Nz + TZ - 1) / TZ;
    src.assign(nty * ntz, {});
    tap.assign(nty * ntz, {});
    for (int s = 0; s < nsrc; ++s) {
        const int c = src_cell[s];
        if (c < 0) continue;
        const int i = c / (Ny * Nz), rem = c % (Ny * Nz);
        const int j = rem / Nz, kk = rem % Nz;
        for (int ty = 0; ty < nty; ++ty) {
            const int e = j - (ty * TYO - 1);
            if (e < 0 || e >= EY) continue;
            for (int tz = 0; tz < ntz; ++tz) {
                const int cc = kk - tz * TZ;
                const bool mainCell = cc >= 0 && cc < TZ;
                const bool haloCell = (cc == -1 || cc == TZ) && e >= 1 && e <= TYO;
                if (mainCell || haloCell) src[ty + nty * tz].push_back(make_int2(ekey(i, e, cc), src_id[s]));
            }
        }
    }
    for (int q = 0; q < ntap; ++q) {
        const int c = tap_cell[q];
        const int i = c / (Ny * Nz), rem = c % (Ny * Nz);
        const int j = rem / Nz, kk = rem % Nz;
        const int ty = j / TYO, tz = kk / TZ;
        tap[ty + nty * tz].push_back(make_int2(ekey(i, j - (ty * TYO - 1), kk - tz * TZ), tap_slot[q]));
    }
}

int f2_rows_for(int tile_y) { return tile_y == 8 ? 1 : 2; }
static int f2_tyo(int rows) { return rows == 1 ? f2::Geo2<1>::TYO : f2::Geo2<2>::TYO; }
int f2_tile_rows(int tile_y) { return f2_tyo(f2_rows_for(tile_y)); }

// Per-tile source and tap tables (uploaded when sources or receivers change).
static int build_tables(AcousticFDTDState* st) {
    using namespace f2;
    const int TYO = st->order == 4 ? f4_tile_rows() : f2_tyo(st->f2_rows);
    const int nt = ((st->Ny + TYO - 1) / TYO) * ((st->Nz + TZ - 1) / TZ);
    std::vector<std::vector<int2>> src, tap;
    std::vector<int> sid(st->npsrc), tslot(8 * (size_t)st->nrec);
    for (int s = 0; s < st->npsrc; ++s) sid[s] = s;
    for (size_t q = 0; q < tslot.size(); ++q) tslot[q] = (int)q;
    if (st->order == 4)
        f4_tile_lists(st->Ny, st->Nz, st->npsrc, st->h_pidx.data(), sid.data(), 8 * st->nrec,
                      st->h_ridx.data(), tslot.data(), src, tap);
    else
        f2_tile_lists(st->Nx, st->Ny, st->Nz, TYO, st->npsrc, st->h_pidx.data(), sid.data(),
                      8 * st->nrec, st->h_ridx.data(), tslot.data(), src, tap);
    auto by_key = [](const int2& x, const int2& y) { return x.x < y.x || (x.x == y.x && x.y < y.y); };
    std::vector<int> ss(nt + 1, 0), ts(nt + 1, 0);
    std::vector<int2> sflat, tflat;
    st->f2_ok = 1;
    for (int t = 0; t < nt; ++t) {
        std::stable_sort(src[t].begin(), src[t].end(), by_key);
        std::stable_sort(tap[t].begin(), tap[t].end(), by_key);
        if ((int)src[t].size() > MAXS || (int)tap[t].size() > MAXT) st->f2_ok = 0;
        ss[t] = (int)sflat.size();
        sflat.insert(sflat.end(), src[t].begin(), src[t].end());
        ts[t] = (int)tflat.size();
        tflat.insert(tflat.end(), tap[t].begin(), tap[t].end());
    }
    ss[nt] = (int)sflat.size();
    ts[nt] = (int)tflat.size();
    if (sflat.empty()) sflat.push_back(make_int2(-1, 0));
    if (tflat.empty()) tflat.push_back(make_int2(-1, 0));
    auto up = [&](void** d, const void* h, size_t bytes) -> hipError_t {
        if (*d) { hipError_t e = hipFree(*d); *d = nullptr; if (e != hipSuccess) return e; }
        hipError_t e = hipMalloc(d, bytes);
        if (e != hipSuccess) return e;
        return hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, st->stream);
    };
    AFDTD_CHECK(up((void**)&st->d_f2_src_start, ss.data(), ss.size() * sizeof(int)));
    AFDTD_CHECK(up((void**)&st->d_f2_src, sflat.data(), sflat.size() * sizeof(int2)));
    AFDTD_CHECK(up((void**)&st->d_f2_tap_start, ts.data(), ts.size() * sizeof(int)));
    AFDTD_CHECK(up((void**)&st->d_f2_tap, tflat.data(), tflat.size() * sizeof(int2)));
    const size_t ntap = st->nrec > 0 ? 8 * (size_t)st->nrec : 8;
    if (!st->d_f2_tapbuf || st->f2_tapcap < ntap) {
        if (st->d_f2_tapbuf) AFDTD_CHECK(hipFree(st->d_f2_tapbuf));
        st->d_f2_tapbuf = nullptr;
        AFDTD_CHECK(hipMalloc(&st->d_f2_tapbuf, ntap * sizeof(float)));
        AFDTD_CHECK(hipMemsetAsync(st->d_f2_tapbuf, 0, ntap * sizeof(float), st->stream));
        st->f2_tapcap = ntap;
    }
    AFDTD_CHECK(hipStreamSynchronize(st->stream));
    st->f2_dirty = 0;
    return FDTD_OK;
}

int fused2_prepare(AcousticFDTDState* st) {
    return st->f2_dirty ? build_tables(st) : FDTD_OK;
}

hipError_t launch_fused2_pair(AcousticFDTDState* st, int step) {
    using namespace f2;
    if (st->order == 4) return launch_fused4_pair(st, step);
    Pair2Args a;
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
    a.seg = st->fused_seg_x > 0 ? st->fused_seg_x : 64;
    if (a.seg > st->Nx) a.seg = st->Nx;
    const int TYO = f2_tyo(st->f2_rows);
    a.nty = (st->Ny + TYO - 1) / TYO;
    a.ntz = (st->Nz + TZ - 1) / TZ;
    a.nseg = (st->Nx + a.seg - 1) / a.seg;
    a.src_start = st->d_f2_src_start;
    a.src = (const int2*)st->d_f2_src;
    a.tap_start = st->d_f2_tap_start;
    a.tap = (const int2*)st->d_f2_tap;
    a.psig = st->d_psig;
    a.plen = st->npsrc > 0 ? st->plen : 0;
    a.step = step;
    a.tapbuf = st->d_f2_tapbuf;
    const int nb = a.nty * a.ntz * a.nseg;
    hipError_t e;
    switch (st->storage) {
        case FDTD_STORE_BF16:
            e = st->hetero ? launch2<bf16_t, true>(st, a, nb) : launch2<bf16_t, false>(st, a, nb);
            break;
        case FDTD_STORE_F16:
            e = st->hetero ? launch2<f16_t, true>(st, a, nb) : launch2<f16_t, false>(st, a, nb);
            break;
        default:
            e = st->hetero ? launch2<float, true>(st, a, nb) : launch2<float, false>(st, a, nb);
    }
    if (e != hipSuccess) return e;
    for (int c = 0; c < 3; ++c) {
        void* t = st->V[c];
        st->V[c] = st->V2[c];
        st->V2[c] = t;
    }
    return hipSuccess;
}

hipError_t launch_fused2_taps(AcousticFDTDState* st, int trace_row) {
    if (trace_row < 0 || st->nrec <= 0 || !st->d_trace) return hipSuccess;
    hipLaunchKernelGGL(k_tap_trace, dim3((st->nrec + 255) / 256), dim3(256), 0, st->stream,
                       st->d_f2_tapbuf, st->d_rwgt, st->nrec, st->d_trace, trace_row,
                       st->d_status);
    return hipGetLastError();
}

hipError_t launch_tap_trace(const float* tapbuf, const float* rwgt, int nrec, float* trace,
                            int trace_row, const int* status, hipStream_t s) {
    if (trace_row < 0 || nrec <= 0 || !trace) return hipSuccess;
    hipLaunchKernelGGL(k_tap_trace, dim3((nrec + 255) / 256), dim3(256), 0, s, tapbuf, rwgt, nrec,
                       trace, trace_row, status);
    return hipGetLastError();
}

namespace {
template <typename T, int R, bool PML>
hipError_t launch_batch2(const F2BatchLaunch& L, const Pair2Args& a) {
    hipLaunchKernelGGL((k_fused2_batch<T, R, PML>), dim3(L.ntiles), dim3(64, f2::Geo2<R>::NW), 0,
                       L.stream, a, L.rooms, L.tile_room);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_batch2_t(const F2BatchLaunch& L, const Pair2Args& a) {
    const bool pml = L.common.n_pml > 0;
    if (L.tile_y == 8) return pml ? launch_batch2<T, 1, true>(L, a) : launch_batch2<T, 1, false>(L, a);
    return pml ? launch_batch2<T, 2, true>(L, a) : launch_batch2<T, 2, false>(L, a);
}
}  // namespace

hipError_t launch_fused2_batch(const F2BatchLaunch& L) {
    Pair2Args a;
    memset(&a, 0, sizeof(a));
    a.p = L.common;
    a.f.Pin = L.Pin;
    a.f.Pout = L.Pout;
    for (int c = 0; c < 3; ++c) {
        a.f.Vin[c] = L.Vin[c];
        a.f.Vout[c] = L.Vout[c];
    }
    a.f.bu = nullptr;
    a.f.kap = nullptr;
    a.f.xi = nullptr;
    a.alpha = L.alpha;
    a.pdamp = L.pdamp;
    a.status = L.status;
    a.seg = L.seg;
    a.src_start = L.src_start;
    a.src = L.src;
    a.tap_start = L.tap_start;
    a.tap = L.tap;
    a.psig = L.psig;
    a.plen = L.plen;
    a.step = L.step;
    a.tapbuf = L.tapbuf;
    switch (L.storage) {
        case FDTD_STORE_BF16: return launch_batch2_t<bf16_t>(L, a);
        case FDTD_STORE_F16: return launch_batch2_t<f16_t>(L, a);
        default: return launch_batch2_t<float>(L, a);
    }
}
