// afdtd_fused4.hip — two leapfrog steps per sweep for the reference's
// order-4 stencil (acoustic_fdtd.cpp:1099-1234; coefficients 9/8, -1/24,
// :889-890; 2nd-order gradient at faces next to the domain edge,
// :1107-1112; zero-padded divergence, :894-925).
//
// The single-step sweep reads and writes every field once per step.  This
// sweep advances n -> n+2 per pass over HBM.  A workgroup owns a tile of
// TY = 24 rows x 64 columns (y, z) and marches an x-segment of it; per
// x-plane q it runs four stages, each trailing the one before:
//   A faces    (q)   V^{n+1/2} at plane q on the extended tile (the tile plus a
//                    three-cell ring: a stage-B face 1.5 cells outside the tile
//                    reads P^{n+1} three cells out) and on the extra face rows
//                    / columns the ring's divergence reads;
//   A pressure (q-1) P^{n+1} on the extended tile (+ point sources of step n+1);
//   B faces    (q-3) V^{n+3/2} of the tile's faces and of the faces up to 1.5
//                    cells outside it, written to HBM for the tile's own faces;
//   B pressure (q-4) P^{n+2} of the tile, written to HBM.
// P^{n+1} and V^{n+1/2} never go to HBM.  P^n is read once per plane (the tile
// plus a six-cell halo), V^{n-1/2} once (the extended faces).
//
// Work mapping (768 threads = 12 wavefronts, one workgroup per CU: 154 KB of
// LDS): lane = column, wave w owns the tile rows 2w and 2w+1 (slots 0 and 1:
// both stages); waves 0..8 also own one ring cell (slot 2: waves 0..5 the
// ring rows -3, -2, -1, 24, 25, 26 at columns 0..63, waves 6..8 the 180
// cells of columns -3..-1 and 64..66).  A slot keeps its x-direction history
// in registers (P^n of planes q-1..q+3, its stage-A x-faces of four planes,
// its stage-B x-faces of four planes), so x-neighbours never go through LDS;
// y / z neighbours do: P^n of the current plane, the stage-A y / z faces and
// P^{n+1} of the last four planes, the stage-B y / z faces of two.  The
// ring's extra stage-A faces (y-face rows -5, -4, 27; z-face columns -5, -4,
// 67) are one more task for waves 0..5; the P^n halo cells outside the
// extended tile are loaded by all threads.  One barrier per plane: every LDS
// value an iteration reads was written before the previous barrier (or by
// the same lane earlier in the iteration).  The loop body is unrolled four
// times so that the register queues rotate by phase instead of moving.
//
// Arithmetic is the single-step sweep's, in the same order, so the result
// equals two single steps (and the oracle) bit for bit.  The PML damping of
// a face or cell is applied only on planes and tiles within reach of the
// absorbing layer (elsewhere the multiplier is exactly 1).  Receiver taps of
// step n+1 go to the tap buffer as in the order-2 sweep; the stability check
// never falls on step n+1 (the host schedules pairs so).
#include "afdtd_fused_tile.cuh"

#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

using fused::FusedPtrs;
using fused::xcd_remap;

namespace {

namespace f4 {
constexpr int TZ = 64;                        // own columns (one per lane)
constexpr int TY = 24;                        // own rows (two per wave)
constexpr int NW = 12;                        // wavefronts
constexpr int NT = 64 * NW;
constexpr int H = 3;                          // P^{n+1} ring (cells)
constexpr int G = 6;                          // P^n halo (cells)
constexpr int ER = TY + 2 * H, EC = TZ + 2 * H;   // extended tile 30 x 70 (rows -3..26, cols -3..66)
constexpr int GR = TY + 2 * G, GC = TZ + 2 * G;   // P^n region 36 x 76 (rows -6..29, cols -6..69)
constexpr int AYR = TY + 9, AZC = TZ + 9;         // stage-A y-face rows -5..27, z-face cols -5..67
constexpr int BYR = TY + 3, BZC = TZ + 3;         // stage-B y-face rows -2..24, z-face cols -2..64
constexpr int PLN = GR * GC, AYN = AYR * EC, AZN = ER * AZC, P1N = ER * EC;
constexpr int BYN = BYR * TZ, BZN = TY * BZC;
constexpr int NRG = GR * GC - ER * EC;            // halo cells outside the extended tile (636)
constexpr int NRGS = (NRG + NT - 1) / NT;         // of them per thread
constexpr int NCR = 6 * ER;                       // column-ring cells (180)
constexpr int NYX = 3 * EC, NZX = 3 * ER;         // extra stage-A y faces (210), z faces (90)
constexpr int MAXS = F2_MAX_SRC;
constexpr int MAXT = F2_MAX_TAP;
__device__ __forceinline__ int gix(int r, int c) { return (r + G) * GC + (c + G); }
__device__ __forceinline__ int p1ix(int e, int c) { return (e + H) * EC + (c + H); }
__device__ __forceinline__ int ayix(int f, int c) { return (f + 5) * EC + (c + H); }
__device__ __forceinline__ int azix(int e, int g) { return (e + H) * AZC + (g + 5); }
__device__ __forceinline__ int byix(int f, int c) { return (f + 2) * TZ + c; }
__device__ __forceinline__ int bzix(int e, int g) { return e * BZC + (g + 2); }
// list key: plane << 12 | (extended row + 3) << 7 | (extended column + 3)
__host__ __device__ __forceinline__ int ekey(int plane, int e, int c) {
    return (plane << 12) | ((e + H) << 7) | (c + H);
}
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             0x00020000);
}
__device__ __forceinline__ float bld(Rsrc r, uint32_t off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void bst(Rsrc r, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, 0, 2);
}
// gradient across a face from its four P taps (2nd order when not inner)
__device__ __forceinline__ float grad4(float pm, float pl, float pr, float pp, bool inner,
                                       float C1, float C2) {
    return inner ? C1 * (pr - pl) + C2 * (pp - pm) : pr - pl;
}
__device__ __forceinline__ uint32_t bld8(Rsrc r, uint32_t off) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}
// Pair code of a cell (GEOM build): bit 0 wall / plate cell (pressure kept),
// bits 1..3 closed x / y / z face, bit 7 plate z face (as the single-step
// code), bit 4 some divergence tap of the cell is fraction-weighted, bit 5
// the cell is above or below an active plate node.
constexpr uint32_t G2_TAPCUT = 16u, G2_PLNB = 32u;
constexpr int PLS = TY * 5;                   // plate ring: 5 terms per tile row and plane
}  // namespace f4

struct Pair4Args {
    StencilParams p;
    FusedPtrs f;
    const float* alpha;
    const float* pdamp;
    const int* status;
    int seg, nty, ntz, nseg;
    const int* src_start;
    const int2* src;        // (key, source index), sorted by key then source index
    const int* tap_start;
    const int2* tap;        // (key, tap slot), sorted by key
    const float* psig;
    int plen;
    int step;               // global index of the first of the two steps
    float* tapbuf;
    const int* items;       // work-item list (nullptr: item = block)
    // cell geometry (GEOM build)
    const uint8_t* g2;      // per-cell pair code (G2_* bits)
    const uint32_t* g2p;    // the pair codes of planes 4g .. 4g+3 of a cell in word g*sx + cell
    const float* tw;        // [3][N] divergence-tap weights: the face fraction where the
                            // tap is weighted, else exactly 1
    const float* vpl;       // plate velocity of step n per plate node (i*Ny + j)
    const float* pre;       // [5][Nplate] step n+1's plate terms: T1, c3, c4, fb, w
    float* pab;             // [Nplate][2] P^{n+1} above / below each active node
    int Nplate, piz, plate_on;
    float pa, inv_dt;       // plate coefficient a, 1/dt
};

template <bool PML, bool GEOM, bool CUT>
__device__ __forceinline__ void fused4_tile(const Pair4Args& a, int ty, int tz, int sg,
                                            int tile) {
    using namespace f4;
    using fused::G_SOLID;
    using fused::G_CLOSED;
    using fused::G_CUT;
    using fused::G_PLATE;
    __shared__ float sPl[GEOM ? 8 * PLS : 1];   // step n+1's plate terms, planes mod 8
    __shared__ float sVp[GEOM ? 2 * ER : 1];    // step n's plate velocity, rows -3..26, planes mod 2
    __shared__ float sP[2 * PLN];
    __shared__ float sAy[4 * AYN];
    __shared__ float sAz[4 * AZN];
    __shared__ float sP1[4 * P1N];
    __shared__ float sBy[2 * BYN];
    __shared__ float sBz[2 * BZN];
    __shared__ float sAlpha[65], sPd[65];
    __shared__ int sSrcKey[MAXS], sTapKey[MAXT], sTapSlot[MAXT];
    __shared__ float sAmp[MAXS];

    const StencilParams& p = a.p;
    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz, sx = Ny * Nz;
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int tid = lane + 64 * w;
    const int j0 = ty * TY, k0 = tz * TZ;
    const int i0 = sg * a.seg, i1 = min(i0 + a.seg, Nx);
    const int npml = PML ? p.n_pml : 0;
    if (PML)
        for (int d = tid; d <= npml && d < 65; d += NT) { sAlpha[d] = a.alpha[d]; sPd[d] = a.pdamp[d]; }
    const float C1 = p.C1, C2 = p.C2, vcoef = p.vel_coeff, kap0 = p.bulk_coeff, dt = p.dt;
    const float* __restrict__ Pin = (const float*)a.f.Pin;
    float* __restrict__ Pout = (float*)a.f.Pout;
    const float* __restrict__ Vxi = (const float*)a.f.Vin[0];
    const float* __restrict__ Vyi = (const float*)a.f.Vin[1];
    const float* __restrict__ Vzi = (const float*)a.f.Vin[2];
    float* __restrict__ Vxo = (float*)a.f.Vout[0];
    float* __restrict__ Vyo = (float*)a.f.Vout[1];
    float* __restrict__ Vzo = (float*)a.f.Vout[2];
    const uint32_t pb = (uint32_t)sx * 4u;                   // plane bytes
    constexpr uint32_t OOB = 0x80000000u;                     // past every plane: load 0, store dropped
    // descriptor of plane pl of field F (empty outside the grid: loads give 0)
    auto rP = [&](const float* F, int pl) __attribute__((always_inline)) {
        return (pl >= 0 && pl < Nx) ? rsrc(F + (size_t)pl * sx, pb) : rsrc(F, 0u);
    };
    // PML depth of one axis position; cells outside the grid are clamped to
    // the table (their values are never used)
    auto pd = [&](int x, int n) __attribute__((always_inline)) { return min(pml_depth(x, n, npml), npml); };
    auto damp = [&](int dL, int dR) __attribute__((always_inline)) { return face_damp(sAlpha, dL, dR, dt); };

    // ── slots: 0, 1 own cells (rows 2w, 2w+1), 2 a ring cell ──
    int se[3], sc[3];
    bool sact[3];
    se[0] = 2 * w; sc[0] = lane; sact[0] = true;
    se[1] = 2 * w + 1; sc[1] = lane; sact[1] = true;
    const bool ringW = w < 9;           // waves with a ring slot
    if (w < 6) {
        se[2] = w < 3 ? w - 3 : TY + w - 3;
        sc[2] = lane;
        sact[2] = true;
    } else {
        const int idx = (w - 6) * 64 + lane, cs = idx % 6;
        se[2] = -3 + idx / 6;
        sc[2] = cs < 3 ? cs - 3 : TZ + cs - 3;
        sact[2] = ringW && idx < NCR;
        if (!sact[2]) { se[2] = 0; sc[2] = 0; }
    }
    uint32_t off[3];             // byte offset of the slot cell in a plane
    bool cellv[3], vyv[3], vzv[3], yin[3], zin[3];
    int dz3[3];                  // PML depths of the slot cell | y face << 8 | z face << 16
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int j = j0 + se[s], k = k0 + sc[s];
        cellv[s] = sact[s] && j >= 0 && j < Ny && k >= 0 && k < Nz;
        off[s] = cellv[s] ? (uint32_t)(j * Nz + k) * 4u : OOB;
        vyv[s] = cellv[s] && j <= Ny - 2;
        vzv[s] = cellv[s] && k <= Nz - 2;
        yin[s] = j > 0 && j < Ny - 2;
        zin[s] = k > 0 && k < Nz - 2;
        const int py = pd(j, Ny), pz = pd(k, Nz);
        dz3[s] = max(py, pz) | (max(pd(j + 1, Ny), pz) << 8) | (max(py, pd(k + 1, Nz)) << 16);
    }
    auto D0 = [&](int s) __attribute__((always_inline)) { return dz3[s] & 0xFF; };
    auto DY = [&](int s) __attribute__((always_inline)) { return (dz3[s] >> 8) & 0xFF; };
    auto DZ = [&](int s) __attribute__((always_inline)) { return dz3[s] >> 16; };
    // stage-B roles of the ring slot: y-faces of rows -2, -1, TY (waves 1..3),
    // z-faces of columns -2, -1, 64 on the tile rows (waves 6..8)
    const bool ringBy = w >= 1 && w <= 3;
    const bool ringBzW = w >= 6 && w < 9;
    const bool ringBz = ringBzW && sact[2] && se[2] >= 0 && se[2] < TY &&
                        (sc[2] == -2 || sc[2] == -1 || sc[2] == TZ);
    // ── misc task: an extra stage-A face (waves 0..3 y, 4..5 z) ──
    int mf = 0, mc = 0;          // y: face row f, column c;  z: row e, face column g
    bool mact = false, mval = false, min_ = false;
    uint32_t moff = OOB;
    int mdL = 0, mdR = 0;
    if (w < 4) {
        const int idx = w * 64 + lane, t = idx / EC;
        mact = idx < NYX;
        mf = t == 0 ? -5 : t == 1 ? -4 : TY + 3;
        mc = -3 + idx % EC;
        const int j = j0 + mf, k = k0 + mc;
        mval = mact && j >= 0 && j <= Ny - 2 && k >= 0 && k < Nz;
        min_ = j > 0 && j < Ny - 2;
        if (mval) moff = (uint32_t)(j * Nz + k) * 4u;
        mdL = max(pd(j, Ny), pd(k, Nz));
        mdR = max(pd(j + 1, Ny), pd(k, Nz));
    } else if (w < 6) {
        const int idx = (w - 4) * 64 + lane, t = idx / ER;
        mact = idx < NZX;
        mc = t == 0 ? -5 : t == 1 ? -4 : TZ + 3;
        mf = -3 + idx % ER;
        const int j = j0 + mf, k = k0 + mc;
        mval = mact && j >= 0 && j < Ny && k >= 0 && k <= Nz - 2;
        min_ = k > 0 && k < Nz - 2;
        if (mval) moff = (uint32_t)(j * Nz + k) * 4u;
        mdL = max(pd(j, Ny), pd(k, Nz));
        mdR = max(pd(j, Ny), pd(k + 1, Nz));
    }
    if (!mact) { mf = 0; mc = 0; }
    // ── P^n halo cells outside the extended tile ──
    uint32_t roff[NRGS];
    int rlds[NRGS];
#pragma unroll
    for (int m = 0; m < NRGS; ++m) {
        const int r = tid + NT * m;
        int row = 0, col = 0;
        if (r < 3 * GC) { row = -G + r / GC; col = -G + r % GC; }
        else if (r < 6 * GC) { row = TY + H + (r - 3 * GC) / GC; col = -G + (r - 3 * GC) % GC; }
        else {
            const int r2 = r - 6 * GC, cs = r2 % 6;
            row = -H + r2 / 6;
            col = cs < 3 ? -G + cs : TZ + H + (cs - 3);
        }
        const int j = j0 + row, k = k0 + col;
        const bool v = r < NRG && j >= 0 && j < Ny && k >= 0 && k < Nz;
        roff[m] = v ? (uint32_t)(j * Nz + k) * 4u : OOB;
        rlds[m] = r < NRG ? gix(row, col) : -1;
    }

    // ── source and tap lists of the tile ──
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

    // ── per-slot pipeline registers ──
    // At phase P (iteration index mod 4) entry u of a face queue (u = 0
    // newest) lives in physical slot (P - u) & 3, and P^n of plane p in slot
    // (p - qf + 1) & 3.  The LDS rings are indexed the same way, by the plane
    // relative to qf (slot (p - qf) & 3 or & 1), so every ring offset is a
    // compile-time constant of the phase.
    float pn[3][4], pnl[3];             // P^n of planes q-1..q+2; q+3 in flight
    float ax[3][4];                     // stage-A x-faces of planes q, q-1, q-2, q-3
    float ovx[3], ovy[3], ovz[3], ovm;  // V^{n-1/2} of the next plane's stage-A faces
    float bx[2][4];                     // own: stage-B x-faces of planes q-3 .. q-6
    float rg[NRGS];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u) pn[s][u] = ax[s][u] = 0.0f;
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int u = 0; u < 4; ++u) bx[s][u] = 0.0f;
    // Damping multipliers of the slot's faces and cell on planes away from
    // the x-sides of the layer (the x depth is 0 there, so they depend on the
    // slot alone; exactly 1 for cells outside the layer).  The CUT build
    // evaluates them per plane from the LDS tables instead (the same values:
    // with x depth 0 the per-plane expression is this one), which frees
    // twelve registers for its weighted-tap path.
    float mX[3], mY[3], mZ[3], mP[3], mM = 1.0f;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        mX[s] = mY[s] = mZ[s] = mP[s] = 1.0f;
        if (PML) {
            mX[s] = face_damp(a.alpha, D0(s), D0(s), dt);
            mY[s] = face_damp(a.alpha, D0(s), DY(s), dt);
            mZ[s] = face_damp(a.alpha, D0(s), DZ(s), dt);
            mP[s] = a.pdamp[D0(s)];
        }
    }
    if (PML) mM = face_damp(a.alpha, mdL, mdR, dt);
    // planes whose faces or cells reach the x-sides of the layer
    auto xpml = [&](int q) __attribute__((always_inline)) {
        return PML && (q < npml + 1 || q >= Nx - npml - 1);
    };

    // ── cell geometry (GEOM) ──
    // gh[s]: pair codes of the slot cell, byte d = plane q-d at iteration q;
    // gn[s] / gmn: the codes of plane q+1 in flight.
    uint64_t gh[3] = {0, 0, 0};
    uint32_t gn[3] = {0, 0, 0}, gmc = 0, gmn = 0;
    uint32_t gw[3] = {0, 0, 0}, gwm = 0;     // packed codes of planes q+1 .. q+4 (phase 0)
    float plv = 0.0f, vpv = 0.0f;
    // Plate loader lanes, for tiles whose z faces reach the plate's two face
    // columns: lanes 0..9 of wave w fetch step n+1's five terms of the rows
    // 2w, 2w+1 (stage B), lanes 10..12 of waves 0..9 step n's plate velocity
    // of the rows -3+3w .. -1+3w (stage A), for the next plane.
    const bool plTile = GEOM && a.plate_on && a.piz >= k0 - 5 && a.piz - 1 <= k0 + TZ + 3;
    const int plRow = 2 * w + lane / 5, plComp = lane % 5;
    const bool plLane = plTile && lane < 10 && j0 + plRow < Ny;
    const int vpRow = -H + 3 * w + (lane - 10);
    const bool vpLane = plTile && w < 10 && lane >= 10 && lane < 13 && j0 + vpRow >= 0 &&
                        j0 + vpRow < Ny;
    // tap weight of face (axis, cell); 0 outside the grid, where the tap is
    // +0 and +0 * 0 = +0 leaves it as it is
    const uint32_t twb = GEOM ? (uint32_t)Nx * (uint32_t)sx * 4u : 0u;
    const Rsrc TW0 = rsrc(a.tw, twb), TW1 = rsrc(a.tw + (GEOM ? (size_t)Nx * sx : 0), twb),
               TW2 = rsrc(a.tw + (GEOM ? 2 * (size_t)Nx * sx : 0), twb);
    auto twt = [&](int ax, int i, int j, int k) __attribute__((always_inline)) {
        const uint32_t o = (i >= 0 && i < Nx && j >= 0 && j < Ny && k >= 0 && k < Nz)
                               ? (uint32_t)(i * sx + j * Nz + k) * 4u : OOB;
        return bld(ax == 0 ? TW0 : ax == 1 ? TW1 : TW2, o);
    };
    // a tap is weighted where its weight is not exactly 1 (the single-step
    // sweep's rule: the fraction bit is clear exactly where the weight is 1)
    auto wtap = [](float v, float wgt) __attribute__((always_inline)) {
        return __float_as_uint(wgt) != 0x3F800000u ? v * wgt : v;
    };
    // divergence of cell (i, j, k) whose taps include weighted faces (cold
    // path); taps x+, x-, x++, x--, then y, then z
    auto cut_dv = [&](int i, int j, int k, float xp, float xm, float xpp, float xmm, float yp,
                      float ym, float ypp, float ymm, float zp, float zm, float zpp,
                      float zmm) __attribute__((always_inline)) {
        const float dx1 = wtap(xp, twt(0, i, j, k)) - wtap(xm, twt(0, i - 1, j, k));
        const float dx2 = wtap(xpp, twt(0, i + 1, j, k)) - wtap(xmm, twt(0, i - 2, j, k));
        const float dy1 = wtap(yp, twt(1, i, j, k)) - wtap(ym, twt(1, i, j - 1, k));
        const float dy2 = wtap(ypp, twt(1, i, j + 1, k)) - wtap(ymm, twt(1, i, j - 2, k));
        const float dz1 = wtap(zp, twt(2, i, j, k)) - wtap(zm, twt(2, i, j, k - 1));
        const float dz2 = wtap(zpp, twt(2, i, j, k + 1)) - wtap(zmm, twt(2, i, j, k - 2));
        return C1 * ((dx1 + dy1) + dz1) + C2 * ((dx2 + dy2) + dz2);
    };
    auto rG = [&](int pl) __attribute__((always_inline)) {
        return (pl >= 0 && pl < Nx) ? rsrc(a.g2 + (size_t)pl * sx, (uint32_t)sx) : rsrc(a.g2, 0u);
    };


    const int qf = i0 - 5;              // first plane of the stage-A faces
    // prologue: P^n of planes qf-1..qf+2 at the slots, plane qf into LDS,
    // V^{n-1/2} of plane qf
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const Rsrc r = rP(Pin, qf - 1 + u);
#pragma unroll
        for (int s = 0; s < 3; ++s) pn[s][u] = bld(r, off[s]);
    }
    {
        const Rsrc r = rP(Pin, qf);
        float* Pw = sP;                 // plane qf: ring slot 0
#pragma unroll
        for (int m = 0; m < NRGS; ++m) {
            const float v = bld(r, roff[m]);
            if (rlds[m] >= 0) Pw[rlds[m]] = v;
        }
#pragma unroll
        for (int s = 0; s < 3; ++s)
            if (sact[s]) Pw[gix(se[s], sc[s])] = pn[s][1];
        const Rsrc X = rP(Vxi, qf), Y = rP(Vyi, qf), Z = rP(Vzi, qf);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            ovx[s] = bld(X, off[s]);
            ovy[s] = bld(Y, off[s]);
            ovz[s] = bld(Z, off[s]);
        }
        ovm = bld(w < 4 ? Y : Z, moff);
        if constexpr (GEOM) {
            const Rsrc g = rG(qf);
#pragma unroll
            for (int s = 0; s < 3; ++s) gh[s] = bld8(g, off[s] >> 2);
            gmc = bld8(g, moff >> 2);
            if (vpLane && qf >= 0 && qf < Nx)
                sVp[(qf & 1) * ER + vpRow + H] = a.vpl[(size_t)qf * Ny + j0 + vpRow];
        }
    }
    __syncthreads();
    // cursors into the cached lists and the key each points at
    int scur = 0, tcur = 0;
    int skey = ns > 0 ? __builtin_amdgcn_readfirstlane(sSrcKey[0]) : 0x7FFFFFFF;
    int tkey = nt > 0 ? __builtin_amdgcn_readfirstlane(sTapKey[0]) : 0x7FFFFFFF;

    auto body = [&](const int q, auto phase) __attribute__((always_inline)) {
        constexpr int P = decltype(phase)::value;
        constexpr int N0 = P & 3, N1 = (P + 3) & 3, N2 = (P + 2) & 3, N3 = (P + 1) & 3;  // newest .. oldest
        constexpr int M0 = N0, M1 = (P + 1) & 3, M2 = (P + 2) & 3, M3 = (P + 3) & 3;      // P^n planes q-1 .. q+2
        const int qa = q - 1, qb = q - 3, qc = q - 4;
        // halo cells of plane q+1 and the slots' P^n of plane q+3
        {
            const Rsrc r1 = rP(Pin, q + 1), r3 = rP(Pin, q + 3);
#pragma unroll
            for (int m = 0; m < NRGS; ++m) rg[m] = bld(r1, roff[m]);
#pragma unroll
            for (int s = 0; s < 3; ++s) pnl[s] = bld(r3, off[s]);
        }
        if constexpr (GEOM) {
            // pair codes of plane q+1 (one load per four planes: the segment
            // start is a multiple of 4, so phase 0 has q+1 = 4g); step n+1's
            // plate terms of plane q+1
            if constexpr (P == 0) {
                const int grp = (q + 1) >> 2;
                const Rsrc g = (grp >= 0 && 4 * grp < Nx)
                                   ? rsrc(a.g2p + (size_t)grp * sx, (uint32_t)sx * 4u)
                                   : rsrc(a.g2p, 0u);
#pragma unroll
                for (int s = 0; s < 3; ++s) gw[s] = __builtin_amdgcn_raw_buffer_load_b32(g, off[s], 0, 0);
                gwm = __builtin_amdgcn_raw_buffer_load_b32(g, moff, 0, 0);
            }
#pragma unroll
            for (int s = 0; s < 3; ++s) gn[s] = (gw[s] >> (8 * P)) & 0xFFu;
            gmn = (gwm >> (8 * P)) & 0xFFu;
            if (plLane && q + 1 >= 0 && q + 1 < Nx)
                plv = a.pre[(size_t)plComp * a.Nplate + (size_t)(q + 1) * Ny + j0 + plRow];
            if (vpLane && q + 1 >= 0 && q + 1 < Nx)
                vpv = a.vpl[(size_t)(q + 1) * Ny + j0 + vpRow];
        }

        // ════ stage-A faces of plane q ════
        {
            const float* Pq = sP + (P & 1) * PLN;
            float* Ayw = sAy + (P & 3) * AYN;
            float* Azw = sAz + (P & 3) * AZN;
            const bool fvx = q >= 0 && q <= Nx - 2, xin = q > 0 && q < Nx - 2;
            const bool qv = q >= 0 && q < Nx;
            const bool xd = xpml(q);
            const int dq = PML ? pd(q, Nx) : 0, dq1 = PML ? pd(q + 1, Nx) : 0;
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                if (s == 2 && !ringW) continue;
                const int e = se[s], c = sc[s];
                const float pm = pn[s][M0], pl = pn[s][M1];
                // x-face between planes q and q+1
                float nv = ovx[s] - vcoef * grad4(pm, pl, pn[s][M2], pn[s][M3], xin, C1, C2);
                if (PML) nv = nv * ((CUT || xd) ? damp(max(dq, D0(s)), max(dq1, D0(s))) : mX[s]);
                const uint32_t g0 = GEOM ? (uint32_t)gh[s] & 0xFFu : 0u;   // plane q
                float nxv = (fvx && cellv[s]) ? nv : 0.0f;
                if (GEOM && (g0 & G_CLOSED)) nxv = 0.0f;
                ax[s][N0] = nxv;
                // y-face between rows e and e+1
                nv = ovy[s] - vcoef * grad4(Pq[gix(e - 1, c)], pl, Pq[gix(e + 1, c)],
                                            Pq[gix(e + 2, c)], yin[s], C1, C2);
                if (PML) nv = nv * ((CUT || xd) ? damp(max(dq, D0(s)), max(dq, DY(s))) : mY[s]);
                float ny = (qv && vyv[s]) ? nv : 0.0f;
                if (GEOM && (g0 & (G_CLOSED << 1))) ny = 0.0f;
                // z-face between columns c and c+1
                nv = ovz[s] - vcoef * grad4(Pq[gix(e, c - 1)], pl, Pq[gix(e, c + 1)],
                                            Pq[gix(e, c + 2)], zin[s], C1, C2);
                if (PML) nv = nv * ((CUT || xd) ? damp(max(dq, D0(s)), max(dq, DZ(s))) : mZ[s]);
                float nz = (qv && vzv[s]) ? nv : 0.0f;
                if constexpr (GEOM) {
                    if (g0 & (G_CLOSED << 2)) nz = 0.0f;
                    if (g0 & G_PLATE) nz = sVp[(q & 1) * ER + e + H];
                }
                if (sact[s]) {
                    Ayw[ayix(e, c)] = ny;
                    Azw[azix(e, c)] = nz;
                }
            }
            if (w < 4) {            // extra y-face (mf | mf+1) at column mc
                float nv = ovm - vcoef * grad4(Pq[gix(mf - 1, mc)], Pq[gix(mf, mc)],
                                               Pq[gix(mf + 1, mc)], Pq[gix(mf + 2, mc)], min_, C1, C2);
                if (PML) nv = nv * (xd ? damp(max(dq, mdL), max(dq, mdR)) : mM);
                nv = (qv && mval) ? nv : 0.0f;
                if (GEOM && (gmc & (G_CLOSED << 1))) nv = 0.0f;
                if (mact) Ayw[ayix(mf, mc)] = nv;
            } else if (w < 6) {     // extra z-face (mc | mc+1) at row mf
                float nv = ovm - vcoef * grad4(Pq[gix(mf, mc - 1)], Pq[gix(mf, mc)],
                                               Pq[gix(mf, mc + 1)], Pq[gix(mf, mc + 2)], min_, C1, C2);
                if (PML) nv = nv * (xd ? damp(max(dq, mdL), max(dq, mdR)) : mM);
                nv = (qv && mval) ? nv : 0.0f;
                if constexpr (GEOM) {
                    if (gmc & (G_CLOSED << 2)) nv = 0.0f;
                    if (gmc & G_PLATE) nv = sVp[(q & 1) * ER + mf + H];
                }
                if (mact) Azw[azix(mf, mc)] = nv;
            }
        }
        // V^{n-1/2} of plane q+1 (consumed next iteration; with geometry after
        // the stage-A pressure, whose weighted-tap path needs the registers)
        auto load_old_v = [&]() __attribute__((always_inline)) {
            const Rsrc X = rP(Vxi, q + 1), Y = rP(Vyi, q + 1), Z = rP(Vzi, q + 1);
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                if (s == 2 && !ringW) continue;
                ovx[s] = bld(X, off[s]);
                ovy[s] = bld(Y, off[s]);
                ovz[s] = bld(Z, off[s]);
            }
            if (w < 6) ovm = bld(w < 4 ? Y : Z, moff);
        };
        if constexpr (!GEOM) load_old_v();

        // ════ stage-A pressure of plane qa = q-1 ════
        if (qa >= i0 - 3) {
            const float* Ayr = sAy + ((P + 3) & 3) * AYN;
            const float* Azr = sAz + ((P + 3) & 3) * AZN;
            float* P1w = sP1 + ((P + 3) & 3) * P1N;
            const bool qav = qa >= 0 && qa < Nx;
            const bool xd = xpml(qa);
            const int dqa = PML ? pd(qa, Nx) : 0;
            float pnew[3];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                pnew[s] = 0.0f;
                if (s == 2 && !ringW) continue;
                const int e = se[s], c = sc[s];
                const float xp = ax[s][N1], xm = ax[s][N2], xpp = ax[s][N0], xmm = ax[s][N3];
                const float yp = Ayr[ayix(e, c)], ym = Ayr[ayix(e - 1, c)];
                const float ypp = Ayr[ayix(e + 1, c)], ymm = Ayr[ayix(e - 2, c)];
                const float zp = Azr[azix(e, c)], zm = Azr[azix(e, c - 1)];
                const float zpp = Azr[azix(e, c + 1)], zmm = Azr[azix(e, c - 2)];
                float dv = C1 * ((xp - xm) + (yp - ym) + (zp - zm))
                         + C2 * ((xpp - xmm) + (ypp - ymm) + (zpp - zmm));
                const uint32_t g1 = GEOM ? (uint32_t)(gh[s] >> 8) & 0xFFu : 0u;   // plane qa
                if (CUT && (g1 & G2_TAPCUT))
                    dv = cut_dv(qa, j0 + e, k0 + c, xp, xm, xpp, xmm, yp, ym, ypp, ymm, zp, zm, zpp, zmm);
                float out = pn[s][M0] - kap0 * dv;
                if (PML) out = out * ((CUT || xd) ? sPd[max(dqa, D0(s))] : mP[s]);
                if (GEOM && (g1 & G_SOLID)) out = pn[s][M0];          // walls keep P
                pnew[s] = (qav && cellv[s]) ? out : 0.0f;
            }
            // point sources of step n+1, in source order
            while ((skey >> 12) <= qa) {
                if ((skey >> 12) == qa) {
                    const int e = ((skey >> 7) & 31) - H, c = (skey & 127) - H;
                    const float amp = sAmp[scur];
#pragma unroll
                    for (int s = 0; s < 3; ++s)
                        if (sact[s] && se[s] == e && sc[s] == c) pnew[s] = pnew[s] + amp;
                }
                ++scur;
                skey = scur < ns ? __builtin_amdgcn_readfirstlane(sSrcKey[scur]) : 0x7FFFFFFF;
            }
#pragma unroll
            for (int s = 0; s < 3; ++s)
                if (sact[s]) P1w[p1ix(se[s], sc[s])] = pnew[s];
            // P^{n+1} beside the plate, for the host-side plate update of step n+1
            if constexpr (GEOM) {
                if (qa >= i0 && qa < i1) {
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        if (((uint32_t)(gh[s] >> 8) & G2_PLNB) && cellv[s]) {
                            const size_t pn_ = (size_t)qa * Ny + j0 + se[s];
                            a.pab[2 * pn_ + (k0 + sc[s] == a.piz + 1 ? 0 : 1)] = pnew[s];
                        }
                    }
                }
            }
            // receiver taps of step n+1 (own planes and cells)
            while ((tkey >> 12) <= qa) {
                if ((tkey >> 12) == qa) {
                    const int e = ((tkey >> 7) & 31) - H, c = (tkey & 127) - H;
                    if (qa >= i0 && qa < i1 && c == lane && (e >> 1) == w)
                        a.tapbuf[sTapSlot[tcur]] = (e & 1) ? pnew[1] : pnew[0];
                }
                ++tcur;
                tkey = tcur < nt ? __builtin_amdgcn_readfirstlane(sTapKey[tcur]) : 0x7FFFFFFF;
            }
        }

        if constexpr (GEOM) load_old_v();

        // ════ stage-B faces of plane qb = q-3 ════
        // P^{n+1} of planes qb-1 .. qb+2 are in LDS slots (q-4)&3 .. (q-1)&3; the
        // slot of plane q-1 was written earlier in this iteration by the same lane
        if (qb >= i0 - 2) {
            const float* P1m = sP1 + (P & 3) * P1N;
            const float* P1b = sP1 + ((P + 1) & 3) * P1N;
            const float* P1r = sP1 + ((P + 2) & 3) * P1N;
            const float* P1p = sP1 + ((P + 3) & 3) * P1N;
            const bool fvb = qb >= 0 && qb <= Nx - 2, xinb = qb > 0 && qb < Nx - 2;
            const bool qbv = qb >= 0 && qb < Nx, ownq = qb >= i0 && qb < i1;
            const bool xd = xpml(qb);
            const int dqb = PML ? pd(qb, Nx) : 0, dqb1 = PML ? pd(qb + 1, Nx) : 0;
            const Rsrc X = rP(Vxo, ownq ? qb : -1), Y = rP(Vyo, ownq ? qb : -1),
                       Z = rP(Vzo, ownq ? qb : -1);
#pragma unroll
            for (int s = 0; s < 2; ++s) {   // x-face between planes qb and qb+1
                const int ix = p1ix(se[s], sc[s]);
                float nv = ax[s][N3] - vcoef * grad4(P1m[ix], P1b[ix], P1r[ix], P1p[ix], xinb, C1, C2);
                if (PML) nv = nv * ((CUT || xd) ? damp(max(dqb, D0(s)), max(dqb1, D0(s))) : mX[s]);
                nv = (fvb && cellv[s]) ? nv : 0.0f;
                if (GEOM && ((uint32_t)(gh[s] >> 24) & G_CLOSED)) nv = 0.0f;
                bx[s][N0] = nv;
                if (ownq) bst(X, off[s], nv);
            }
            if (ownq) {
                const float* Ayo = sAy + ((P + 1) & 3) * AYN;    // stage-A faces of plane qb
                const float* Azo = sAz + ((P + 1) & 3) * AZN;
                float* Byw = sBy + ((P + 1) & 1) * BYN;
                float* Bzw = sBz + ((P + 1) & 1) * BZN;
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    if (s == 2 && !ringBy) continue;
                    const int e = se[s], c = sc[s];
                    float nv = Ayo[ayix(e, c)] - vcoef * grad4(P1b[p1ix(e - 1, c)], P1b[p1ix(e, c)],
                                                               P1b[p1ix(e + 1, c)], P1b[p1ix(e + 2, c)],
                                                               yin[s], C1, C2);
                    if (PML) nv = nv * ((CUT || xd) ? damp(max(dqb, D0(s)), max(dqb, DY(s))) : mY[s]);
                    nv = (qbv && vyv[s]) ? nv : 0.0f;
                    if (GEOM && ((uint32_t)(gh[s] >> 24) & (G_CLOSED << 1))) nv = 0.0f;
                    Byw[byix(e, c)] = nv;
                    if (s < 2) bst(Y, off[s], nv);
                }
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    if (s == 2 && !ringBzW) continue;
                    // ring lanes without a stage-B z-face read a harmless cell
                    const int e = (s < 2 || ringBz) ? se[s] : 0, c = (s < 2 || ringBz) ? sc[s] : 0;
                    float nv = Azo[azix(e, c)] - vcoef * grad4(P1b[p1ix(e, c - 1)], P1b[p1ix(e, c)],
                                                               P1b[p1ix(e, c + 1)], P1b[p1ix(e, c + 2)],
                                                               zin[s], C1, C2);
                    if (PML) nv = nv * ((CUT || xd) ? damp(max(dqb, D0(s)), max(dqb, DZ(s))) : mZ[s]);
                    nv = (qbv && vzv[s]) ? nv : 0.0f;
                    if constexpr (GEOM) {
                        const uint32_t g3 = (s < 2 || ringBz) ? (uint32_t)(gh[s] >> 24) & 0xFFu : 0u;
                        if (g3 & (G_CLOSED << 2)) nv = 0.0f;
                        if (g3 & G_PLATE) {
                            // step n+1's plate update of node (qb, j0+e), as the
                            // plate kernel evaluates it, from P^{n+1} beside the plate
                            const float* pl = sPl + (qb & 7) * PLS + e * 5;
                            const float Pa = a.piz + 1 < Nz ? P1b[p1ix(e, a.piz + 1 - k0)] : 0.0f;
                            const float Pb = a.piz - 1 >= 0 ? P1b[p1ix(e, a.piz - 1 - k0)] : 0.0f;
                            const float wn = pl[0] + a.pa * ((Pb - Pa) + pl[3]) - pl[1] - pl[2];
                            nv = (wn - pl[4]) * a.inv_dt;
                        }
                    }
                    if (s < 2) {
                        Bzw[bzix(e, c)] = nv;
                        bst(Z, off[s], nv);
                    } else if (ringBz) {
                        Bzw[bzix(e, c)] = nv;
                    }
                }
            }
        }

        // ════ stage-B pressure of plane qc = q-4 ════
        if (qc >= i0 && qc < i1) {
            const float* Byr = sBy + (P & 1) * BYN;
            const float* Bzr = sBz + (P & 1) * BZN;
            const float* P1c = sP1 + (P & 3) * P1N;
            const bool xd = xpml(qc);
            const int dqc = PML ? pd(qc, Nx) : 0;
            const Rsrc R = rP(Pout, qc);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int e = se[s], c = sc[s];
                const float xp = bx[s][N1], xm = bx[s][N2], xpp = bx[s][N0], xmm = bx[s][N3];
                const float yp = Byr[byix(e, c)], ym = Byr[byix(e - 1, c)];
                const float ypp = Byr[byix(e + 1, c)], ymm = Byr[byix(e - 2, c)];
                const float zp = Bzr[bzix(e, c)], zm = Bzr[bzix(e, c - 1)];
                const float zpp = Bzr[bzix(e, c + 1)], zmm = Bzr[bzix(e, c - 2)];
                float dv = C1 * ((xp - xm) + (yp - ym) + (zp - zm))
                         + C2 * ((xpp - xmm) + (ypp - ymm) + (zpp - zmm));
                const uint32_t g4 = GEOM ? (uint32_t)(gh[s] >> 32) & 0xFFu : 0u;   // plane qc
                if (CUT && (g4 & G2_TAPCUT))
                    dv = cut_dv(qc, j0 + e, k0 + c, xp, xm, xpp, xmm, yp, ym, ypp, ymm, zp, zm, zpp, zmm);
                float out = P1c[p1ix(e, c)] - kap0 * dv;
                if (PML) out = out * ((CUT || xd) ? sPd[max(dqc, D0(s))] : mP[s]);
                if (!GEOM || !(g4 & G_SOLID)) bst(R, off[s], out);   // walls keep P
            }
        }

        // ── P^n plane q+1 into LDS; the slots' P^n of plane q+3 into the queue ──
        {
            float* Pw = sP + ((P + 1) & 1) * PLN;
#pragma unroll
            for (int m = 0; m < NRGS; ++m)
                if (rlds[m] >= 0) Pw[rlds[m]] = rg[m];
#pragma unroll
            for (int s = 0; s < 3; ++s)
                if (sact[s]) Pw[gix(se[s], sc[s])] = pn[s][M2];
        }
#pragma unroll
        for (int s = 0; s < 3; ++s) pn[s][M0] = pnl[s];
        if constexpr (GEOM) {
            // codes of plane q+1 into the history; the plate rows of plane q+1
            // into the LDS rings
#pragma unroll
            for (int s = 0; s < 3; ++s) gh[s] = (gh[s] << 8) | gn[s];
            gmc = gmn;
            if (plLane) sPl[((q + 1) & 7) * PLS + plRow * 5 + plComp] = plv;
            if (vpLane) sVp[((q + 1) & 1) * ER + vpRow + H] = vpv;
        }
        __syncthreads();
    };

    const int qend = i1 + 3;
#pragma unroll 1
    for (int q = qf; q <= qend; q += 4) {
        body(q, std::integral_constant<int, 0>());
        if (q + 1 <= qend) body(q + 1, std::integral_constant<int, 1>());
        if (q + 2 <= qend) body(q + 2, std::integral_constant<int, 2>());
        if (q + 3 <= qend) body(q + 3, std::integral_constant<int, 3>());
    }
}

// CUT: the tiles whose cells have fraction-weighted divergence taps (the
// weighted path costs registers, so the other tiles run without it)
template <bool PML, bool GEOM, bool CUT>
__global__ void __launch_bounds__(f4::NT) k_fused4(Pair4Args a) {
    if (*a.status) return;
    // each XCD takes a contiguous range of work items; within a segment the
    // items run in bands of up to four z-tiles, z fastest, that sweep y
    int t = xcd_remap(blockIdx.x, gridDim.x);
    if (a.items) t = a.items[t];
    const int per_sg = a.nty * a.ntz;
    const int sg = t / per_sg;
    t -= sg * per_sg;
    const int band = t / (a.nty * 4);
    const int gz = min(4, a.ntz - 4 * band);
    t -= band * a.nty * 4;
    const int ty = t / gz;
    const int tz = 4 * band + t % gz;
    fused4_tile<PML, GEOM, CUT>(a, ty, tz, sg, ty + a.nty * tz);
}

// Pair codes (bit layout at G2_TAPCUT) from the single-step codes: the
// wall, closed and plate bits of the cell, and whether any of its twelve
// divergence taps (x faces i, i-1, i+1, i-2 of the cell's line, likewise y
// and z) carries the fraction bit of its axis.
__global__ void __launch_bounds__(256)
k_geo2_codes(int Nx, int Ny, int Nz, const uint8_t* __restrict__ code, uint8_t* __restrict__ g2) {
    const size_t N = (size_t)Nx * Ny * Nz;
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const int k = (int)(c % Nz);
    const size_t t = c / Nz;
    const int j = (int)(t % Ny), i = (int)(t / Ny);
    const size_t sx = (size_t)Ny * Nz;
    using namespace fused;
    uint32_t out = code[c] & (G_SOLID | (G_CLOSED * 7u) | G_PLATE);
    bool cut = false;
    const int d[4] = {0, -1, 1, -2};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (i + d[u] >= 0 && i + d[u] < Nx) cut |= (code[c + d[u] * (long)sx] & G_CUT) != 0;
        if (j + d[u] >= 0 && j + d[u] < Ny) cut |= (code[c + d[u] * (long)Nz] & (G_CUT << 1)) != 0;
        if (k + d[u] >= 0 && k + d[u] < Nz) cut |= (code[c + d[u]] & (G_CUT << 2)) != 0;
    }
    if (cut) out |= f4::G2_TAPCUT;
    g2[c] = (uint8_t)out;
}

// Plate-neighbour bit on the cells above and below each active node.
__global__ void __launch_bounds__(256)
k_geo2_plate(int nact, const int* __restrict__ above, const int* __restrict__ below,
             uint8_t* __restrict__ g2) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nact) return;
    if (above[n] >= 0) g2[above[n]] = (uint8_t)(g2[above[n]] | f4::G2_PLNB);
    if (below[n] >= 0) g2[below[n]] = (uint8_t)(g2[below[n]] | f4::G2_PLNB);
}

// Pair codes packed four planes to a word: word g*sx + c holds the codes
// of cell c in planes 4g .. 4g+3 (byte p - 4g; 0 past the grid).
__global__ void __launch_bounds__(256)
k_geo2_pack(int Nx, size_t sx, const uint8_t* __restrict__ g2, uint32_t* __restrict__ g2p) {
    const size_t ng = (size_t)((Nx + 3) / 4);
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= ng * sx) return;
    const size_t g = t / sx, c = t % sx;
    uint32_t w = 0;
    for (int b = 0; b < 4; ++b) {
        const size_t pl = 4 * g + b;
        if (pl < (size_t)Nx) w |= (uint32_t)g2[pl * sx + c] << (8 * b);
    }
    g2p[t] = w;
}

// Divergence-tap weights: the face fraction where the single-step code
// weights the tap, exactly 1 elsewhere (the fraction bit is set only where
// the fraction is not exactly 1).
__global__ void __launch_bounds__(256)
k_tap_weights(size_t N, const uint8_t* __restrict__ code, const float* __restrict__ fx,
              const float* __restrict__ fy, const float* __restrict__ fz, float* __restrict__ tw) {
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const uint32_t g = code[c];
    tw[c] = (g & fused::G_CUT) ? fx[c] : 1.0f;
    tw[N + c] = (g & (fused::G_CUT << 1)) ? fy[c] : 1.0f;
    tw[2 * N + c] = (g & (fused::G_CUT << 2)) ? fz[c] : 1.0f;
}

// Step n+1's plate update split at the pressure term: per active node p,
// with w = w^{n+1}, wp = w^n, the terms T1 = w + vd (w - wp), c3 = bhc bh,
// c4 = bk (bh - bhp), fb and w itself, so that the update is
// ((T1 + a ((Pb - Pa) + fb)) - c3) - c4, the plate kernel's expression
// (acoustic_fdtd.cpp:984-1053) with its 13-tap sums in tap order.
__global__ void __launch_bounds__(256)
k_plate_terms(int nact, const int* __restrict__ act, const int* __restrict__ b13,
              const float* __restrict__ w, const float* __restrict__ wp,
              const float* __restrict__ ext, float vd, float bhc, float bk, int Np,
              float* __restrict__ pre, const int* status) {
    if (status[0]) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nact) return;
    const float W13[13] = {20.f, -8.f, -8.f, -8.f, -8.f, 2.f, 2.f, 2.f, 2.f, 1.f, 1.f, 1.f, 1.f};
    const int p = act[n];
    const int* nb = b13 + 13 * n;
    float bh = 0.0f, bhp = 0.0f;
    for (int t = 0; t < 13; ++t) bh += W13[t] * w[nb[t]];
    for (int t = 0; t < 13; ++t) bhp += W13[t] * wp[nb[t]];
    const float vp = w[p] - wp[p];
    pre[p] = w[p] + vd * vp;
    pre[(size_t)Np + p] = bhc * bh;
    pre[2 * (size_t)Np + p] = bk * (bh - bhp);
    pre[3 * (size_t)Np + p] = ext[p];
    pre[4 * (size_t)Np + p] = w[p];
}

constexpr int kPairFail = 9;    // status word while a paired step's plate guard is handled

// Step n+1's plate update from the terms and the sweep's P^{n+1} beside the
// plate; the guard (acoustic_fdtd.cpp:1032-1040) marks the pair for the
// single-step redo of step n.
__global__ void __launch_bounds__(256)
k_plate_pair(int nact, const int* __restrict__ act, const int* __restrict__ above,
             const int* __restrict__ below, const float* __restrict__ pre, int Np,
             const float* __restrict__ pab, float a, float* __restrict__ wt, int* status,
             int step_index) {
    if (status[0]) return;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nact) return;
    const int p = act[n];
    const float Pa = above[n] >= 0 ? pab[2 * (size_t)p] : 0.0f;
    const float Pb = below[n] >= 0 ? pab[2 * (size_t)p + 1] : 0.0f;
    const float wn = pre[p] + a * ((Pb - Pa) + pre[3 * (size_t)Np + p]) - pre[(size_t)Np + p]
                   - pre[2 * (size_t)Np + p];
    if (!isfinite(wn) || fabsf(wn) > 0.02f) {
        if (atomicCAS(&status[0], 0, kPairFail) == 0) status[1] = step_index;
        return;
    }
    wt[p] = wn;
}

// After the redo of step n: its point sources (in source order, as the
// post-step kernel adds them), then the plate failure status of step n+1.
__global__ void k_pair_redo_end(float* __restrict__ P, int npsrc, const int* __restrict__ pidx,
                                const float* __restrict__ psig, int plen, int step_index,
                                int* status) {
    if (status[0] != kPairFail) return;
    if (step_index < plen)
        for (int s = 0; s < npsrc; ++s) P[pidx[s]] = P[pidx[s]] + psig[(size_t)s * plen + step_index];
    status[0] = 2;
}

}  // namespace

bool fused4_geo_supported(const AcousticFDTDState* st) {
    return st->order == 4 && st->storage == FDTD_STORE_F32 && !st->hetero && st->n_pml <= 64 &&
           fused_geo_supported(st);
}

hipError_t launch_geo_pair_codes(AcousticFDTDState* st) {
    ++st->g2_gen;                   // the item lists follow the codes
    if (!st->d_g2) {
        hipError_t e = hipMalloc(&st->d_g2, (size_t)st->N);
        if (e != hipSuccess) { st->d_g2 = nullptr; return e; }
    }
    hipLaunchKernelGGL(k_geo2_codes, dim3((unsigned)(((size_t)st->N + 255) / 256)), dim3(256), 0,
                       st->stream, st->Nx, st->Ny, st->Nz, st->d_geo, st->d_g2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (st->has_frac) {
        if (!st->d_tw) {
            e = hipMalloc(&st->d_tw, 3 * sizeof(float) * (size_t)st->N);
            if (e != hipSuccess) { st->d_tw = nullptr; return e; }
        }
        hipLaunchKernelGGL(k_tap_weights, dim3((unsigned)(((size_t)st->N + 255) / 256)), dim3(256),
                           0, st->stream, (size_t)st->N, st->d_geo, st->d_fx, st->d_fy, st->d_fz,
                           st->d_tw);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (st->nact > 0) {
        hipLaunchKernelGGL(k_geo2_plate, dim3((st->nact + 255) / 256), dim3(256), 0, st->stream,
                           st->nact, st->d_above, st->d_below, st->d_g2);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const size_t sx = (size_t)st->Ny * st->Nz, npk = (size_t)((st->Nx + 3) / 4) * sx;
    if (!st->d_g2p) {
        e = hipMalloc(&st->d_g2p, sizeof(uint32_t) * npk);
        if (e != hipSuccess) { st->d_g2p = nullptr; return e; }
    }
    hipLaunchKernelGGL(k_geo2_pack, dim3((unsigned)((npk + 255) / 256)), dim3(256), 0, st->stream,
                       st->Nx, sx, st->d_g2, st->d_g2p);
    e = hipGetLastError();
    if (e != hipSuccess || st->nact <= 0) return e;
    if (!st->d_ppre) {
        e = hipMalloc(&st->d_ppre, 5 * sizeof(float) * st->Nplate);
        if (e != hipSuccess) { st->d_ppre = nullptr; return e; }
        e = hipMemsetAsync(st->d_ppre, 0, 5 * sizeof(float) * st->Nplate, st->stream);
        if (e != hipSuccess) return e;
    }
    if (!st->d_pab) {
        e = hipMalloc(&st->d_pab, 2 * sizeof(float) * st->Nplate);
        if (e != hipSuccess) { st->d_pab = nullptr; return e; }
        e = hipMemsetAsync(st->d_pab, 0, 2 * sizeof(float) * st->Nplate, st->stream);
    }
    return e;
}

bool fused4_supported(const AcousticFDTDState* st) {
    return st->order == 4 && fused_supported(st) && st->storage == FDTD_STORE_F32 && !st->hetero;
}

int f4_tile_rows() { return f4::TY; }

// Per-tile source and tap lists: a source goes to every tile whose extended
// tile (rows -3..TY+2, columns -3..66) holds its cell, a tap to the tile
// that owns its cell.
void f4_tile_lists(int Ny, int Nz, int nsrc, const int* src_cell, const int* src_id, int ntap,
                   const int* tap_cell, const int* tap_slot, std::vector<std::vector<int2>>& src,
                   std::vector<std::vector<int2>>& tap) {
    using namespace f4;
    const int nty = (Ny + TY - 1) / TY, ntz = (Nz + TZ - 1) / TZ;
    src.assign(nty * ntz, {});
    tap.assign(nty * ntz, {});
    for (int s = 0; s < nsrc; ++s) {
        const int c = src_cell[s];
        if (c < 0) continue;
        const int i = c / (Ny * Nz), rem = c % (Ny * Nz);
        const int j = rem / Nz, kk = rem % Nz;
        for (int ty = 0; ty < nty; ++ty) {
            const int e = j - ty * TY;
            if (e < -H || e >= TY + H) continue;
            for (int tz = 0; tz < ntz; ++tz) {
                const int cc = kk - tz * TZ;
                if (cc < -H || cc >= TZ + H) continue;
                src[ty + nty * tz].push_back(make_int2(ekey(i, e, cc), src_id[s]));
            }
        }
    }
    for (int q = 0; q < ntap; ++q) {
        const int c = tap_cell[q];
        const int i = c / (Ny * Nz), rem = c % (Ny * Nz);
        const int j = rem / Nz, kk = rem % Nz;
        const int ty = j / TY, tz = kk / TZ;
        tap[ty + nty * tz].push_back(make_int2(ekey(i, j - ty * TY, kk - tz * TZ), tap_slot[q]));
    }
}

static Pair4Args pair4_args(AcousticFDTDState* st, int step) {
    using namespace f4;
    Pair4Args a;
    memset(&a, 0, sizeof(a));
    a.p = st->sp;
    a.f.Pin = st->P[st->cur];
    a.f.Pout = st->P[st->cur ^ 1];
    for (int c = 0; c < 3; ++c) {
        a.f.Vin[c] = st->V[c];
        a.f.Vout[c] = st->V2[c];
    }
    a.alpha = st->d_alpha;
    a.pdamp = st->d_pdamp;
    a.status = st->d_status;
    a.seg = st->fused_seg_x > 0 ? st->fused_seg_x : 64;
    if (a.seg > st->Nx) a.seg = st->Nx;
    a.nty = (st->Ny + TY - 1) / TY;
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
    return a;
}

#define AFDTD_RET(expr)                         \
    do {                                        \
        hipError_t _e = (expr);                 \
        if (_e != hipSuccess) return _e;        \
    } while (0)

static void swap_velocities(AcousticFDTDState* st) {
    for (int c = 0; c < 3; ++c) {
        void* t = st->V[c];
        st->V[c] = st->V2[c];
        st->V2[c] = t;
    }
}

hipError_t launch_fused4_pair(AcousticFDTDState* st, int step) {
    using namespace f4;
    const Pair4Args a = pair4_args(st, step);
    const int nb = a.nty * a.ntz * a.nseg;
    if (st->n_pml > 0)
        hipLaunchKernelGGL((k_fused4<true, false, false>), dim3(nb), dim3(64, NW), 0, st->stream, a);
    else
        hipLaunchKernelGGL((k_fused4<false, false, false>), dim3(nb), dim3(64, NW), 0, st->stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    swap_velocities(st);
    return hipSuccess;
}

// Per work item (x-segment sg of tile (ty, tz), numbered as k_fused4 numbers
// them): does any cell the item updates (stage A: the extended tile over
// planes i0-3 .. i1+2) carry the weighted-tap bit?
__global__ void __launch_bounds__(256)
k_item_cut(const uint8_t* __restrict__ g2, int Nx, int Ny, int Nz, int seg, int nty, int ntz,
           int* __restrict__ flag) {
    using namespace f4;
    int t = blockIdx.x;
    const int per_sg = nty * ntz;
    const int sg = t / per_sg;
    t -= sg * per_sg;
    const int band = t / (nty * 4);
    const int gz = min(4, ntz - 4 * band);
    t -= band * nty * 4;
    const int ty = t / gz, tz = 4 * band + t % gz;
    const int i0 = max(sg * seg - 3, 0), i1 = min(sg * seg + seg + 3, Nx);
    const int ja = max(ty * TY - H, 0), jb = min(ty * TY + TY + H, Ny);
    const int ka = max(tz * TZ - H, 0), kb = min(tz * TZ + TZ + H, Nz);
    const int nk = kb - ka, nj = jb - ja;
    int any = 0;
    for (int i = i0; i < i1 && !any; ++i)
        for (int r = threadIdx.x; r < nj * nk; r += 256) {
            const int j = ja + r / nk, k = ka + r % nk;
            if (g2[((size_t)i * Ny + j) * Nz + k] & G2_TAPCUT) { any = 1; break; }
        }
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) flag[blockIdx.x] = any;
}

// Work-item lists of the geometry sweep (rebuilt when the codes or the
// segment length change): items without weighted taps first, then those
// with them, each in item order.
static hipError_t geo_item_lists(AcousticFDTDState* st, const Pair4Args& a) {
    if (st->d_g2_items && st->g2_items_seg == a.seg && st->g2_items_gen == st->g2_gen)
        return hipSuccess;
    const int nb = a.nty * a.ntz * a.nseg;
    std::vector<int> flag(nb, 0), list;
    if (st->has_frac) {
        int* d_flag = nullptr;
        AFDTD_RET(hipMalloc(&d_flag, sizeof(int) * nb));
        hipLaunchKernelGGL(k_item_cut, dim3(nb), dim3(256), 0, st->stream, st->d_g2, st->Nx,
                           st->Ny, st->Nz, a.seg, a.nty, a.ntz, d_flag);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(flag.data(), d_flag, sizeof(int) * nb, hipMemcpyDeviceToHost,
                               st->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
        (void)hipFree(d_flag);
        AFDTD_RET(e);
    }
    for (int pass = 0; pass < 2; ++pass)
        for (int t = 0; t < nb; ++t)
            if ((flag[t] != 0) == (pass == 1)) list.push_back(t);
    st->g2_nslow = 0;
    for (int t = 0; t < nb; ++t) st->g2_nslow += flag[t] != 0;
    st->g2_nfast = nb - st->g2_nslow;
    if (st->d_g2_items) AFDTD_RET(hipFree(st->d_g2_items));
    st->d_g2_items = nullptr;
    AFDTD_RET(hipMalloc(&st->d_g2_items, sizeof(int) * nb));
    AFDTD_RET(hipMemcpyAsync(st->d_g2_items, list.data(), sizeof(int) * nb, hipMemcpyHostToDevice,
                             st->stream));
    AFDTD_RET(hipStreamSynchronize(st->stream));
    st->g2_items_seg = a.seg;
    st->g2_items_gen = st->g2_gen;
    return hipSuccess;
}

hipError_t launch_fused4_geo_pair(AcousticFDTDState* st, int step, int trace_row0) {
    using namespace f4;
    if (!st->d_geo || !st->d_g2 || (st->nact > 0 && (!st->d_ppre || !st->d_pab || !st->d_vpl)))
        return hipErrorInvalidValue;
    const int nact = st->nact;
    const unsigned pb = (unsigned)((nact + 255) / 256);
    if (nact > 0) {
        hipLaunchKernelGGL(k_plate_terms, dim3(pb), dim3(256), 0, st->stream, nact, st->d_act,
                           st->d_b13, st->d_w[0], st->d_w[1], st->d_ext, st->plate_vd,
                           st->plate_bh, st->plate_bk, st->Nplate, st->d_ppre, st->d_status);
        AFDTD_RET(hipGetLastError());
    }
    Pair4Args a = pair4_args(st, step);
    a.g2 = st->d_g2;
    a.g2p = st->d_g2p;
    a.tw = st->d_tw;
    a.vpl = st->d_vpl;
    a.pre = st->d_ppre;
    a.pab = st->d_pab;
    a.Nplate = st->Nplate;
    a.piz = st->plate_iz;
    a.plate_on = nact > 0;
    a.pa = st->plate_a;
    a.inv_dt = 1.0f / st->dt;
    // 24-plane segments by default: more work items per launch than the
    // plain pair's 64 (395 x 549 x 148 body, ms per step at segments of
    // 16 / 24 / 32 / 40 / 48 / 64 planes: 0.432 / 0.418 / 0.431 / 0.424 /
    // 0.448 / 0.472); a segment is a whole number of 4-plane code words
    a.seg = st->fused_seg_x > 0 ? st->fused_seg_x : 24;
    a.seg = (a.seg + 3) & ~3;
    if (a.seg >= st->Nx) a.seg = st->Nx;
    a.nseg = (st->Nx + a.seg - 1) / a.seg;
    AFDTD_RET(geo_item_lists(st, a));
    // The tiles with weighted taps run on a second stream beside the others
    // (disjoint outputs, shared read-only inputs), so that neither launch
    // leaves compute units idle at its tail.
    const bool side = st->g2_nslow > 0 && st->g2_nfast > 0;
    if (side) {
        if (!st->g2_stream) {
            AFDTD_RET(hipStreamCreateWithFlags(&st->g2_stream, hipStreamNonBlocking));
            AFDTD_RET(hipEventCreateWithFlags(&st->g2_ev[0], hipEventDisableTiming));
            AFDTD_RET(hipEventCreateWithFlags(&st->g2_ev[1], hipEventDisableTiming));
        }
        AFDTD_RET(hipEventRecord(st->g2_ev[0], st->stream));
        AFDTD_RET(hipStreamWaitEvent(st->g2_stream, st->g2_ev[0], 0));
    }
    hipStream_t slow_stream = side ? st->g2_stream : st->stream;
    a.items = st->d_g2_items + st->g2_nfast;
    if (st->g2_nslow > 0) {
        if (st->n_pml > 0)
            hipLaunchKernelGGL((k_fused4<true, true, true>), dim3(st->g2_nslow), dim3(64, NW), 0,
                               slow_stream, a);
        else
            hipLaunchKernelGGL((k_fused4<false, true, true>), dim3(st->g2_nslow), dim3(64, NW), 0,
                               slow_stream, a);
        AFDTD_RET(hipGetLastError());
    }
    a.items = st->d_g2_items;
    if (st->g2_nfast > 0) {
        if (st->n_pml > 0)
            hipLaunchKernelGGL((k_fused4<true, true, false>), dim3(st->g2_nfast), dim3(64, NW), 0,
                               st->stream, a);
        else
            hipLaunchKernelGGL((k_fused4<false, true, false>), dim3(st->g2_nfast), dim3(64, NW), 0,
                               st->stream, a);
        AFDTD_RET(hipGetLastError());
    }
    if (side) {
        AFDTD_RET(hipEventRecord(st->g2_ev[1], st->g2_stream));
        AFDTD_RET(hipStreamWaitEvent(st->stream, st->g2_ev[1], 0));
    }
    AFDTD_RET(launch_fused2_taps(st, trace_row0));
    if (nact > 0) {
        hipLaunchKernelGGL(k_plate_pair, dim3(pb), dim3(256), 0, st->stream, nact, st->d_act,
                           st->d_above, st->d_below, st->d_ppre, st->Nplate, st->d_pab,
                           st->plate_a, st->d_w[2], st->d_status, step + 1);
        AFDTD_RET(hipGetLastError());
        float* t = st->d_w[2]; st->d_w[2] = st->d_w[1]; st->d_w[1] = st->d_w[0]; st->d_w[0] = t;
        // redo of step n over the sweep's output when step n+1's guard tripped
        // (inputs intact: the sweep writes the other buffers)
        // (the single-step launcher swaps the velocity buffers: swap back)
        AFDTD_RET(launch_fused_step_if(st, kPairFail));
        swap_velocities(st);
        hipLaunchKernelGGL(k_pair_redo_end, dim3(1), dim3(1), 0, st->stream,
                           (float*)st->P[st->cur ^ 1], st->npsrc, st->d_pidx, st->d_psig, st->plen,
                           step, st->d_status);
        AFDTD_RET(hipGetLastError());
    }
    swap_velocities(st);
    return hipSuccess;
}
