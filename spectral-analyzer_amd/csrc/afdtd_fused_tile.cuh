// afdtd_fused_tile.cuh — the fused plane-marching tile update (device code),
// shared by the single-room sweep (afdtd_fused.hip) and the batched-rooms
// sweep (afdtd_batch.hip).  See afdtd_fused.hip for the algorithm.
//
// Work mapping inside a tile of TY rows x 64 columns (NW wavefronts):
//   lane = z column, wave w owns rows w*R .. w*R+R-1 (R = TY/NW);
//   halo face rows of Vy are one extra task each for waves 0..FLO+FHI-1, the
//   halo face columns of Vz are one extra task for the last lanes of the
//   block, so the halo work lands on different wavefronts;
//   every global offset, validity flag and LDS address is computed once per
//   workgroup, so a plane costs only loads, LDS traffic and the stencil flops.
#pragma once
#include "afdtd_common.cuh"

namespace fused {

constexpr int TZ = 64;

struct FusedPtrs {
    const void* Pin;
    void* Pout;
    const void* Vin[3];
    void* Vout[3];
    const float* bu;      // per-cell buoyancy 1/rho (hetero)
    const float* kap;     // per-cell K dt/dx (hetero)
    const float* xi;      // per-cell wall sums (hetero + impedance, edge cells)
};

// Cell geometry of the general sweep (GEOM): one code byte per cell c.
//   bit 0       c is a wall or plate cell: its pressure is not updated;
//   bits 1..3   the x / y / z face (c, c + s_axis) is closed: its velocity is
//               set to 0 after the update (face fraction 0);
//   bits 4..6   the face's fraction is not 1: its divergence tap is V * frac;
//   bit 7       the z face is a plate face: its velocity is set to the
//               plate's velocity (w - w_prev) / dt of this step after the
//               closed-face rule.
// Bits are set only for faces inside the room.
enum : uint32_t { G_SOLID = 1u, G_CLOSED = 2u, G_CUT = 16u, G_PLATE = 128u };

struct GeoPtrs {
    const uint8_t* code;        // N code bytes (GEOM)
    const float* frac[3];       // face fractions, cell-indexed
    const float* vpl;           // plate velocity per plate node (i*Ny + j), this step
    uint32_t ncells;            // bound of the code array (bytes)
};

template <int ORDER>
struct Geo {
    static constexpr int PH = ORDER == 4 ? 3 : 1;   // P halo (cells)
    static constexpr int FLO = ORDER == 4 ? 2 : 1;  // faces below own range
    static constexpr int FHI = ORDER == 4 ? 1 : 0;  // faces above own range
    static constexpr int GHI = ORDER == 4 ? 2 : 1;  // gradient taps above face
    static constexpr int LA = FHI + GHI;            // P-plane lookahead
    static constexpr int NS = LA + 2;               // ring depth (+1: the slot being
                                                    // filled was last read two planes ago)
    static constexpr int NQ = FLO + FHI + 1;        // Vx queue length
};

__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int x = b & 7, per = nb >> 3, rem = nb & 7;
    const int base = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    return base + (b >> 3);
}

// Face update shared by the three components: gradient from the four P taps
// (2nd order at edge faces or for the order-2 stencil,
// acoustic_fdtd.cpp:1107-1112), then (V - coef*g) (:1113).
template <int ORDER>
__device__ __forceinline__ float face_new(float vold, float pm, float pl, float pr, float pp,
                                          bool inner, float coef, float C1, float C2) {
    float g;
    if (ORDER == 4 && inner)
        g = C1 * (pr - pl) + C2 * (pp - pm);
    else
        g = pr - pl;
    return vold - coef * g;
}

template <int ORDER, typename T, bool HET, int TY, int NW, bool GEOM = false>
__device__ __forceinline__ void fused_tile(const StencilParams& p, const FusedPtrs& fp,
                                           const float* __restrict__ alpha_tab,
                                           const float* __restrict__ pdamp_tab,
                                           int ty, int tz, int sg, int seg,
                                           const GeoPtrs& gp = GeoPtrs{}) {
    typedef Geo<ORDER> G;
    typedef Store<T> S;
    static_assert(!GEOM || S::exact, "cell geometry runs with FP32 storage");
    constexpr int PH = G::PH, FLO = G::FLO, FHI = G::FHI, NS = G::NS, LA = G::LA, NQ = G::NQ;
    constexpr int R = TY / NW;                      // rows per thread
    constexpr int PR = TY + 2 * PH, PC = TZ + 2 * PH;
    constexpr int FR = TY + FLO + FHI;              // Vy face rows held in LDS
    constexpr int FC = TZ + FLO + FHI;              // Vz face cols held in LDS
    constexpr int NT = 64 * NW;
    constexpr int HR = 2 * PH;                      // halo rows of a P plane
    constexpr int NHRW = (HR + NW - 1) / NW;        // halo rows per wave
    constexpr int HC = PR * 2 * PH;                 // halo-column elements
    constexpr int NHC = (HC + NT - 1) / NT;         // halo-column elements per thread
    constexpr int NHY = FLO + FHI;                  // halo Vy face rows (one per wave)
    constexpr int NHZ = TY * (FLO + FHI);           // halo Vz faces (last threads)
    static_assert(NHY <= NW, "one halo Vy row per wave");
    static_assert(NHZ <= NT, "halo Vz faces fit the block");
    constexpr int PLANE = PR * PC;

    __shared__ float sP[NS * PLANE];
    __shared__ float sB[HET ? NS * PLANE : 1];
    __shared__ float sVy[FR * TZ];
    __shared__ float sVz[TY * FC];
    __shared__ float sAlpha[65], sPd[65];

    const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
    const int sx = Ny * Nz;
    // blockDim.x is one wavefront, so threadIdx.y is wave-uniform; reading it
    // through readfirstlane lets every per-wave condition be a scalar branch.
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int tid = lane + 64 * w;
    const int j0 = ty * TY, k0 = tz * TZ;
    const int i0 = sg * seg, i1 = min(i0 + seg, Nx);
    const int k = k0 + lane;
    const bool kval = k < Nz;
    const T* __restrict__ Pin = (const T*)fp.Pin;
    T* __restrict__ Pout = (T*)fp.Pout;
    const T* __restrict__ Vxi = (const T*)fp.Vin[0];
    const T* __restrict__ Vyi = (const T*)fp.Vin[1];
    const T* __restrict__ Vzi = (const T*)fp.Vin[2];
    T* __restrict__ Vxo = (T*)fp.Vout[0];
    T* __restrict__ Vyo = (T*)fp.Vout[1];
    T* __restrict__ Vzo = (T*)fp.Vout[2];
    const int npml = p.n_pml;
    if (npml > 0)
        for (int d = tid; d <= npml && d < 65; d += NT) { sAlpha[d] = alpha_tab[d]; sPd[d] = pdamp_tab[d]; }
    const float C1 = p.C1, C2 = p.C2, vcoef = p.vel_coeff, dtdx = p.dtdx, dt = p.dt;
    const float isv = p.inv_sV, sv = p.sV;
    // block floating point: per-row exponents after the N mantissas of each
    // field; exponent index (plane*Ny + j)*nrz + k/64
    constexpr bool BLK = IsBlock<T>::value;
    const int NC = Nx * Ny * Nz, nrz = (Nz + 63) >> 6, esx = Ny * nrz;

    // ── per-thread geometry, computed once ──
    int ownOff[R];          // in-plane offset of own cell (row w*R+r, col lane)
    int ownE[R];            // in-plane exponent index of the own row (block storage)
    bool ownV[R];           // own cell inside the room
    bool vyV[R], vzV[R];    // own Vy / Vz face exists (j <= Ny-2 / k <= Nz-2)
    bool vyIn[R];           // order-4 taps for the own Vy face
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + w * R + r;
        ownOff[r] = j * Nz + k;
        ownV[r] = kval && j < Ny;
        ownE[r] = j < Ny ? j * nrz + tz : 0;
        vyV[r] = kval && j <= Ny - 2;
        vzV[r] = j < Ny && k <= Nz - 2;
        vyIn[r] = j > 0 && j < Ny - 2;
    }
    const bool vzIn = k > 0 && k < Nz - 2;
    // halo rows of the P plane (cols PH .. PH+63)
    int hrOff[NHRW], hrLds[NHRW], hrE[NHRW];
    bool hrV[NHRW];
#pragma unroll
    for (int m = 0; m < NHRW; ++m) {
        const int h = w + NW * m;
        const int row = h < PH ? h : TY + h;
        const int gj = j0 - PH + row;
        hrLds[m] = row * PC + PH + lane;
        hrOff[m] = gj * Nz + k;
        hrV[m] = h < HR && gj >= 0 && gj < Ny && kval;
        hrE[m] = hrV[m] ? gj * nrz + tz : 0;
        if (h >= HR) hrLds[m] = -1;
    }
    // halo columns of the P plane (all PR rows)
    int hcOff[NHC], hcLds[NHC], hcE[NHC];
    bool hcV[NHC];
#pragma unroll
    for (int m = 0; m < NHC; ++m) {
        const int e = tid + NT * m;
        const int row = e / (2 * PH), c = e - row * (2 * PH);
        const int col = c < PH ? c : TZ + c;
        const int gj = j0 - PH + row, gk = k0 - PH + col;
        hcLds[m] = e < HC ? row * PC + col : -1;
        hcOff[m] = gj * Nz + gk;
        hcV[m] = e < HC && gj >= 0 && gj < Ny && gk >= 0 && gk < Nz;
        hcE[m] = hcV[m] ? gj * nrz + (gk >> 6) : 0;
    }
    // halo Vy face row handled by this wave (task A)
    const bool tA = w < NHY;
    const int frA = w < FLO ? w : TY + w;            // LDS face row
    const int jA = j0 - FLO + frA;                   // global face row
    const bool vA = tA && kval && jA >= 0 && jA <= Ny - 2;
    const bool inA = jA > 0 && jA < Ny - 2;
    const int offA = jA * Nz + k;
    const int eA = vA ? jA * nrz + tz : 0;
    // halo Vz face handled by this thread (task B)
    const int uB = tid - (NT - NHZ);
    const bool tB = uB >= 0;
    const int lyB = tB ? uB / (FLO + FHI) : 0;
    const int hcB = tB ? uB - lyB * (FLO + FHI) : 0;
    const int fcB = hcB < FLO ? hcB : TZ + hcB;      // LDS face col
    const int jB = j0 + lyB, kB = k0 - FLO + fcB;
    const bool vB = tB && jB < Ny && kB >= 0 && kB <= Nz - 2;
    const bool inB = kB > 0 && kB < Nz - 2;
    const int offB = jB * Nz + kB;
    const int eB = vB ? jB * nrz + (kB >> 6) : 0;
    // PML depths (used only when npml > 0)
    int dJ[R], dJ1[R];
    const int dK = pml_depth(k, Nz, npml), dK1 = pml_depth(k + 1, Nz, npml);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        dJ[r] = pml_depth(j0 + w * R + r, Ny, npml);
        dJ1[r] = pml_depth(j0 + w * R + r + 1, Ny, npml);
    }
    const int dJA = pml_depth(jA, Ny, npml), dJA1 = pml_depth(jA + 1, Ny, npml);
    const int dJB = pml_depth(jB, Ny, npml), dKB = pml_depth(kB, Nz, npml),
              dKB1 = pml_depth(kB + 1, Nz, npml);
    // own cells on a y/z domain side (impedance walls)
    bool edgeYZ[R], rowEdge[R];                   // rowEdge: the wave's row r has
                                                  // an edge cell (wave-uniform)
    const bool tileZEdge = k0 == 0 || k0 + TZ > Nz - 1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = j0 + w * R + r;
        edgeYZ[r] = (j == 0) | (j == Ny - 1) | (k == 0) | (k == Nz - 1);
        rowEdge[r] = (j == 0) | (j == Ny - 1) | tileZEdge;
    }
    // per-cell wall sums are read only for rooms with per-cell media
    const bool loadXi = HET && p.impedance;
    const int ownLds = (PH + w * R) * PC + PH + lane;    // + r*PC

    // ── staging registers (filled one plane ahead) ──
    // They hold the loaded values unconverted and unmasked: the conversion
    // happens where the value is used, one plane later, so the loads stay in
    // flight across a whole plane of work.  Lanes outside the room load
    // element 0 instead (always inside the field); every face computed from
    // such a value is invalid and replaced by 0, and the P and rho values of
    // cells outside the room are never read by a valid face.
    typedef typename S::Raw RawT;
    RawT stP[R + NHRW + NHC];
    float stB[HET ? R + NHRW + NHC : 1];
    int stE[R + NHRW + NHC];                // row exponents (block storage only)
    // Old velocities / K / wall sums of one plane.  Two sets alternate
    // between even and odd planes (the loop is unrolled by two), so no
    // register copy at the loop latch forces an early wait on the loads.
    struct Old {
        RawT x[R], y[R], z[R], a, b;
        float k[HET ? R : 1], xi[HET ? R : 1];
        int ex[R], ey[R], ez[R], ea, eb;        // row exponents (block storage only)
        // geometry codes (GEOM): of the own Vx face's cell (plane f), of the
        // own cells (plane q), of the halo faces A and B; 0 outside the room
        uint32_t gx[GEOM ? R : 1], gq[GEOM ? R : 1], ga, gb;
    };
    // Fractions and plate displacements of the faces of one plane whose codes
    // ask for them, loaded at the start of the plane (GEOM).
    struct Lazy {
        float fx[R], fy[R], fz[R], fa, fb;
        float vp[R], vpb;
    };
    const __amdgpu_buffer_rsrc_t grs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint8_t*>(gp.code), (short)0, (int)gp.ncells, 0x00020000);
    // code byte at cell offset off, or 0 when !v (the offset lies beyond the
    // descriptor's bound, so the load returns 0)
    auto ldg = [&](bool v, int off) -> uint32_t {
        if constexpr (GEOM)
            return __builtin_amdgcn_raw_buffer_load_b8(grs, v ? (uint32_t)off : 0x80000000u, 0, 0);
        else return 0u;
    };
    Old oE, oO;
    float vq[NQ][R];
#pragma unroll
    for (int s = 0; s < NQ; ++s)
#pragma unroll
        for (int r = 0; r < R; ++r) vq[s][r] = 0.0f;

    auto ldr = [&](const T* __restrict__ F, bool v, int off) { return S::ldr(F, v ? off : 0); };
    auto ldf = [&](const float* __restrict__ F, bool v, int off) { return F[v ? off : 0]; };
    // row exponent of field F at plane pl (block storage; 0 otherwise); the
    // _u form is for rows shared by the whole wavefront (lane 0 always lies
    // inside the room), whose index is made uniform so it is a scalar load
    auto lde = [&](const T* __restrict__ F, int pl, bool v, int eoff) -> int {
        if constexpr (BLK) return bfp_exps(F, NC)[v ? pl * esx + eoff : 0];
        else return 0;
    };
    auto lde_u = [&](const T* __restrict__ F, int pl, bool v, int eoff) -> int {
        if constexpr (BLK)
            return bfp_exps(F, NC)[__builtin_amdgcn_readfirstlane(v ? pl * esx + eoff : 0)];
        else return 0;
    };
    auto dec = [&](RawT x, int e) -> float {
        if constexpr (BLK) return bfp_dec(x, e);
        else return S::cvt(x);
    };
    auto vel = [&](RawT x, int e) {
        const float f = dec(x, e);
        return (S::exact || BLK) ? f : f * isv;
    };
    // store the own row r of a field at plane pl (velocities carry the
    // storage scale sv, pressures none); with block storage every lane of the
    // wavefront takes part in the row's exponent (the caller's condition is
    // wave-uniform), lane 0 writes it
    // e_old: the segment's exponent before this step (wave-uniform), or a
    // value no lane can match (-1000) to force the full reduction
    auto put_row = [&](T* __restrict__ F, int pl, int r, float v, bool valid, bool isV,
                       int e_old) {
        if constexpr (BLK) {
            const int e = bfp_row_exp(v, valid, e_old);
            if (valid)
                __builtin_nontemporal_store(__builtin_bit_cast(uint16_t, blk_enc<RawT>(v, e)),
                                            (uint16_t*)F + pl * sx + ownOff[r]);
            if (lane == 0 && j0 + w * R + r < Ny)
                bfp_exps((void*)F, NC)[pl * esx + ownE[r]] = e;
        } else {
            if (valid) S::st_nt(F, pl * sx + ownOff[r], (S::exact || !isV) ? v : v * sv);
        }
    };
    auto load_plane = [&](int pl) {
        const bool pv = pl >= 0 && pl < Nx;
        const int base = pl * sx;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool v = pv && ownV[r];
            stP[r] = ldr(Pin, v, base + ownOff[r]);
            if (HET) stB[r] = ldf(fp.bu, v, base + ownOff[r]);
            if (BLK) stE[r] = lde_u(Pin, pl, v, ownE[r]);
        }
#pragma unroll
        for (int m = 0; m < NHRW; ++m) {
            const bool v = pv && hrV[m];
            stP[R + m] = ldr(Pin, v, base + hrOff[m]);
            if (HET) stB[R + m] = ldf(fp.bu, v, base + hrOff[m]);
            if (BLK) stE[R + m] = lde_u(Pin, pl, v, hrE[m]);
        }
#pragma unroll
        for (int m = 0; m < NHC; ++m) {
            const bool v = pv && hcV[m];
            stP[R + NHRW + m] = ldr(Pin, v, base + hcOff[m]);
            if (HET) stB[R + NHRW + m] = ldf(fp.bu, v, base + hcOff[m]);
            if (BLK) stE[R + NHRW + m] = lde(Pin, pl, v, hcE[m]);
        }
    };
    auto store_plane = [&](int slot) {
        float* dP = sP + slot * PLANE;
        float* dB = sB + (HET ? slot * PLANE : 0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            dP[ownLds + r * PC] = dec(stP[r], BLK ? stE[r] : 0);
            if (HET) dB[ownLds + r * PC] = stB[r];
        }
#pragma unroll
        for (int m = 0; m < NHRW; ++m)
            if (hrLds[m] >= 0) {
                dP[hrLds[m]] = dec(stP[R + m], BLK ? stE[R + m] : 0);
                if (HET) dB[hrLds[m]] = stB[R + m];
            }
#pragma unroll
        for (int m = 0; m < NHC; ++m)
            if (hcLds[m] >= 0) {
                dP[hcLds[m]] = dec(stP[R + NHRW + m], BLK ? stE[R + NHRW + m] : 0);
                if (HET) dB[hcLds[m]] = stB[R + NHRW + m];
            }
    };
    // Old values of plane q (Vy, Vz, K, wall sums) and of Vx face f.
    auto load_old = [&](int q, int f, Old& o) {
        const bool fv = f >= 0 && f <= Nx - 2;
        const bool qv = q >= i0 && q < i1;
        const int bf = f * sx, bq = q * sx;
        const bool xe = (q == 0) | (q == Nx - 1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            o.x[r] = ldr(Vxi, fv && ownV[r], bf + ownOff[r]);
            o.y[r] = ldr(Vyi, qv && vyV[r], bq + ownOff[r]);
            o.z[r] = ldr(Vzi, qv && vzV[r], bq + ownOff[r]);
            if (BLK) {
                o.ex[r] = lde_u(Vxi, f, fv && ownV[r], ownE[r]);
                o.ey[r] = lde_u(Vyi, q, qv && vyV[r], ownE[r]);
                o.ez[r] = lde_u(Vzi, q, qv && vzV[r], ownE[r]);
            }
            if (HET) o.k[r] = ldf(fp.kap, qv && ownV[r], bq + ownOff[r]);
            if (HET && loadXi && qv && (xe || rowEdge[r]))
                o.xi[r] = ldf(fp.xi, ownV[r] && (xe || edgeYZ[r]), bq + ownOff[r]);
        }
        o.a = ldr(Vyi, qv && vA, bq + offA);
        o.b = ldr(Vzi, qv && vB, bq + offB);
        if (BLK) {
            o.ea = lde_u(Vyi, q, qv && vA, eA);
            o.eb = lde(Vzi, q, qv && vB, eB);
        }
    };
    // Geometry codes of plane q and of the Vx face f (GEOM).  Issued before
    // the plane prefetch, so they have arrived when the next plane starts
    // and the fraction and plate loads they ask for can be issued at once.
    auto load_codes = [&](int q, int f, Old& o) {
        if constexpr (GEOM) {
            const bool fv = f >= 0 && f <= Nx - 2;
            const bool qv = q >= i0 && q < i1;
            const int bf = f * sx, bq = q * sx;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                o.gx[r] = ldg(fv && ownV[r], bf + ownOff[r]);
                o.gq[r] = ldg(qv && ownV[r], bq + ownOff[r]);
            }
            o.ga = ldg(qv && vA, bq + offA);
            o.gb = ldg(qv && vB, bq + offB);
        }
    };
    // The fractions and plate displacements the codes of plane q ask for.
    auto lazy_load = [&](int q, const Old& o, Lazy& z) {
        if constexpr (GEOM) {
            const int bf = (q + FHI) * sx, bq = q * sx, pq = q * Ny;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (o.gx[r] & G_CUT) z.fx[r] = gp.frac[0][bf + ownOff[r]];
                if (o.gq[r] & (G_CUT << 1)) z.fy[r] = gp.frac[1][bq + ownOff[r]];
                if (o.gq[r] & (G_CUT << 2)) z.fz[r] = gp.frac[2][bq + ownOff[r]];
                if (o.gq[r] & G_PLATE) z.vp[r] = gp.vpl[pq + j0 + w * R + r];
            }
            if (o.ga & (G_CUT << 1)) z.fa = gp.frac[1][bq + offA];
            if (o.gb & (G_CUT << 2)) z.fb = gp.frac[2][bq + offB];
            if (o.gb & G_PLATE) z.vpb = gp.vpl[pq + jB];
        }
    };
    auto slot_of = [](int pl) { return ((pl % NS) + NS) % NS; };
    auto slot_add = [](int s0, int d) { const int t = s0 + d; return t >= NS ? t - NS : t; };
    auto damp = [&](int dL, int dR) { return face_damp(sAlpha, dL, dR, dt); };

    // One plane q (ring slot sq0): consumes the old values in `o`, prefetches
    // plane q+1's into `n`.
    auto plane = [&](int q, int sq0, Old& o, Old& n) __attribute__((always_inline)) {
        // The slot of plane q+LA was last read at plane q-2 (ring depth LA+2)
        // and the face tiles at plane q-1 before the previous plane's last
        // barrier, so no barrier is needed before refilling them.
        store_plane(slot_add(sq0, LA));
        Lazy z;
        lazy_load(q, o, z);
        load_codes(q + 1, q + 1 + FHI, n);
        // Issued on every plane, also the last one (whose prefetch is never
        // used: plane q+1 >= i1 is invalid, so its loads read element 0 or a
        // face inside the room).  A conditional prefetch would make the
        // compiler wait for it at once, on the path that skips it.
        load_plane(q + 1 + LA);
        load_old(q + 1, q + 1 + FHI, n);
        __syncthreads();                                   // ring plane visible

        // ── Vx face f = q + FHI of own cells ──
        {
            const int f = q + FHI;
            const bool fv = f >= 0 && f <= Nx - 2;
            const bool inner = f > 0 && f < Nx - 2;
            const bool own = f >= i0 && f < i1;
            // planes f-1 .. f+2 are q+FHI-1 .. q+FHI+2, all >= q-1
            const int sf = slot_add(sq0, FHI);
            const int sf1 = slot_add(sf, 1);
            const int sfm = ORDER == 4 ? sq0 : sf;       // plane f-1 = q when FHI = 1
            const int sf2 = ORDER == 4 ? slot_add(sf, 2) : sf;
            const float* s0 = sP + sf * PLANE;
            const float* s1 = sP + sf1 * PLANE;
            const float* sm = sP + sfm * PLANE;
            const float* s2 = sP + sf2 * PLANE;
            int di0 = 0, di1 = 0;
            if (npml > 0) { di0 = pml_depth(f, Nx, npml); di1 = pml_depth(f + 1, Nx, npml); }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const bool val = fv && ownV[r];
                const int a = ownLds + r * PC;
                float coef = vcoef;
                if (HET) {
                    const float* b0 = sB + sf * PLANE;
                    const float* b1 = sB + sf1 * PLANE;
                    coef = (0.5f * (b0[a] + b1[a])) * dtdx;
                }
                float nv = face_new<ORDER>(vel(o.x[r], BLK ? o.ex[r] : 0), ORDER == 4 ? sm[a] : 0.0f, s0[a], s1[a],
                                           ORDER == 4 ? s2[a] : 0.0f, inner, coef, C1, C2);
                if (npml > 0) nv = nv * damp(imax3(di0, dJ[r], dK), imax3(di1, dJ[r], dK));
                nv = val ? nv : 0.0f;
                float tap = nv;                         // the face's divergence tap
                if constexpr (GEOM) {
                    if (o.gx[r] & G_CLOSED) nv = 0.0f;
                    tap = (o.gx[r] & G_CUT) ? nv * z.fx[r] : nv;
                }
                if (own) put_row(Vxo, f, r, nv, val, true, BLK ? o.ex[r] : 0);
#pragma unroll
                for (int s = 0; s < NQ - 1; ++s) vq[s][r] = vq[s + 1][r];
                vq[NQ - 1][r] = tap;
            }
        }

        if (q >= i0) {
            const float* sq = sP + sq0 * PLANE;
            const float* bq = sB + (HET ? sq0 * PLANE : 0);
            const int dI = pml_depth(q, Nx, npml);
            // ── Vy faces of own rows (face j between rows j and j+1) ──
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int a = ownLds + r * PC;
                const float coef = HET ? (0.5f * (bq[a] + bq[a + PC])) * dtdx : vcoef;
                float nv = face_new<ORDER>(vel(o.y[r], BLK ? o.ey[r] : 0), ORDER == 4 ? sq[a - PC] : 0.0f, sq[a],
                                           sq[a + PC], ORDER == 4 ? sq[a + 2 * PC] : 0.0f,
                                           vyIn[r], coef, C1, C2);
                if (npml > 0) nv = nv * damp(imax3(dI, dJ[r], dK), imax3(dI, dJ1[r], dK));
                nv = vyV[r] ? nv : 0.0f;
                float tap = nv;
                if constexpr (GEOM) {
                    if (o.gq[r] & (G_CLOSED << 1)) nv = 0.0f;
                    tap = (o.gq[r] & (G_CUT << 1)) ? nv * z.fy[r] : nv;
                }
                put_row(Vyo, q, r, nv, vyV[r], true, BLK ? o.ey[r] : 0);
                sVy[(FLO + w * R + r) * TZ + lane] = tap;
            }
            // ── halo Vy face row (task A) ──
            if (tA) {
                const int a = (frA - FLO + PH) * PC + PH + lane;
                const float coef = HET ? (0.5f * (bq[a] + bq[a + PC])) * dtdx : vcoef;
                float nv = face_new<ORDER>(vel(o.a, BLK ? o.ea : 0), ORDER == 4 ? sq[a - PC] : 0.0f, sq[a],
                                           sq[a + PC], ORDER == 4 ? sq[a + 2 * PC] : 0.0f, inA,
                                           coef, C1, C2);
                if (npml > 0) nv = nv * damp(imax3(dI, dJA, dK), imax3(dI, dJA1, dK));
                nv = vA ? nv : 0.0f;
                if constexpr (GEOM) {
                    if (o.ga & (G_CLOSED << 1)) nv = 0.0f;
                    if (o.ga & (G_CUT << 1)) nv = nv * z.fa;
                }
                sVy[frA * TZ + lane] = nv;
            }
            // ── Vz faces of own rows (face k between cols k and k+1) ──
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int a = ownLds + r * PC;
                const float coef = HET ? (0.5f * (bq[a] + bq[a + 1])) * dtdx : vcoef;
                float nv = face_new<ORDER>(vel(o.z[r], BLK ? o.ez[r] : 0), ORDER == 4 ? sq[a - 1] : 0.0f, sq[a],
                                           sq[a + 1], ORDER == 4 ? sq[a + 2] : 0.0f, vzIn, coef,
                                           C1, C2);
                if (npml > 0) nv = nv * damp(imax3(dI, dJ[r], dK), imax3(dI, dJ[r], dK1));
                nv = vzV[r] ? nv : 0.0f;
                float tap = nv;
                if constexpr (GEOM) {
                    if (o.gq[r] & (G_CLOSED << 2)) nv = 0.0f;
                    if (o.gq[r] & G_PLATE) nv = z.vp[r];
                    tap = (o.gq[r] & (G_CUT << 2)) ? nv * z.fz[r] : nv;
                }
                put_row(Vzo, q, r, nv, vzV[r], true, BLK ? o.ez[r] : 0);
                sVz[(w * R + r) * FC + FLO + lane] = tap;
            }
            // ── halo Vz face (task B) ──
            if (tB) {
                const int a = (PH + lyB) * PC + fcB - FLO + PH;
                const float coef = HET ? (0.5f * (bq[a] + bq[a + 1])) * dtdx : vcoef;
                float nv = face_new<ORDER>(vel(o.b, BLK ? o.eb : 0), ORDER == 4 ? sq[a - 1] : 0.0f, sq[a], sq[a + 1],
                                           ORDER == 4 ? sq[a + 2] : 0.0f, inB, coef, C1, C2);
                if (npml > 0) nv = nv * damp(imax3(dI, dJB, dKB), imax3(dI, dJB, dKB1));
                nv = vB ? nv : 0.0f;
                if constexpr (GEOM) {
                    if (o.gb & (G_CLOSED << 2)) nv = 0.0f;
                    if (o.gb & G_PLATE) nv = z.vpb;
                    if (o.gb & (G_CUT << 2)) nv = nv * z.fb;
                }
                sVz[lyB * FC + fcB] = nv;
            }
            __syncthreads();                               // faces visible
            // ── pressure of own cells at plane q ──
            const bool xedge = (q == 0) | (q == Nx - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int fy = (FLO + w * R + r) * TZ + lane;
                const int fz = (w * R + r) * FC + FLO + lane;
                const float xp = vq[FLO][r], xm = vq[FLO - 1][r];
                const float yp = sVy[fy], ym = sVy[fy - TZ];
                const float zp = sVz[fz], zm = sVz[fz - 1];
                float dv;
                if constexpr (ORDER == 2) {
                    dv = (xp - xm) + (yp - ym) + (zp - zm);
                } else {
                    const float xpp = vq[FLO + 1][r], xmm = vq[FLO - 2][r];
                    const float ypp = sVy[fy + TZ], ymm = sVy[fy - 2 * TZ];
                    const float zpp = sVz[fz + 1], zmm = sVz[fz - 2];
                    dv = C1 * ((xp - xm) + (yp - ym) + (zp - zm))
                       + C2 * ((xpp - xmm) + (ypp - ymm) + (zpp - zmm));
                }
                const float pc = sq[ownLds + r * PC];
                const float kap = HET ? o.k[r] : p.bulk_coeff;
                float out;
                if (p.impedance && ownV[r] && (xedge || edgeYZ[r])) {
                    float xi;
                    if (HET) {
                        xi = o.xi[r];
                    } else {
                        const int j = j0 + w * R + r;
                        xi = wall_xi(p.xterm, q, j, k, Nx, Ny, Nz);
                    }
                    out = impedance_update(pc, dv, kap, xi);
                } else {
                    out = pc - kap * dv;
                }
                if (npml > 0) out = out * sPd[imax3(dI, dJ[r], dK)];
                bool pst = ownV[r];
                if constexpr (GEOM) pst = pst && !(o.gq[r] & G_SOLID);   // walls keep P
                put_row(Pout, q, r, out, pst, false, -1000);
            }
        }
    };

    const int q0 = i0 - FLO - FHI;
    for (int pl = q0; pl < q0 + LA; ++pl) {
        load_plane(pl);
        store_plane(slot_of(pl));
    }
    load_codes(q0, q0 + FHI, oE);
    load_plane(q0 + LA);
    load_old(q0, q0 + FHI, oE);
    __syncthreads();

    int sq0 = slot_of(q0);                                 // slot of plane q
    for (int q = q0; q < i1; q += 2) {
        plane(q, sq0, oE, oO);
        sq0 = slot_add(sq0, 1);
        if (q + 1 >= i1) break;
        plane(q + 1, sq0, oO, oE);
        sq0 = slot_add(sq0, 1);
    }
}

}  // namespace fused
