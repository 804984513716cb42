// afdtd_internal.h — device state and kernel launchers of libafdtd.so.
//
// Device layout (HBM): the four fields P, Vx, Vy, Vz share ONE padded layout
// of N = Nx*Ny*Nz elements indexed by the cell index c = k + Nz*(j + Ny*i)
// (z fastest, as the reference's IDX3, acoustic_fdtd.cpp:24-27):
//   Vx[c] = face (i+1/2, j, k), valid for i <= Nx-2 (slot i = Nx-1 stays 0)
//   Vy[c] = face (i, j+1/2, k), valid for j <= Ny-2
//   Vz[c] = face (i, j, k+1/2), valid for k <= Nz-2
// so every stencil tap of every field is a fixed offset from c (+-1, +-Nz,
// +-Ny*Nz) and one z-run of 64 cells is one coalesced 256-byte wavefront
// access for all four fields.  The reference's packed staggered layouts
// (acoustic_fdtd.h:430-437) are produced only at the API boundary, by
// strided device-to-host copies.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/afdtd.h"

#define AFDTD_CHECK(expr)                                                   \
    do {                                                                    \
        hipError_t _e = (expr);                                             \
        if (_e != hipSuccess) {                                             \
            afdtd_set_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
            return FDTD_ERR_DEVICE;                                         \
        }                                                                   \
    } while (0)

void afdtd_set_error(const std::string& msg);

// Per-step constants shared by every stencil kernel (passed by value).
struct StencilParams {
    int Nx, Ny, Nz;
    int order;          // 2 or 4
    int n_pml;          // PML depth (0 = none)
    int impedance;      // locally reacting domain walls (order 2)
    int hetero;         // per-cell rho / kappa arrays present
    int has_solid;      // cell-type array present (walls / plate cells)
    int has_closed;     // face-fraction arrays present (closed-face test)
    int has_frac;       // divergence weighted by face fractions
    float dt, dx, vel_coeff, bulk_coeff;
    float C1, C2;
    float gside[6];     // (1-R)/(1+R) per domain side x-,x+,y-,y+,z-,z+
    float xterm[6];     // homogeneous wall terms (kap/2) * g_side / sqrt(rho K)
    float rho_air, K0;  // homogeneous medium
    float dtdx;         // dt/dx (heterogeneous face coefficient)
    float sV, inv_sV;   // power-of-two velocity storage scale (16-bit storage)
};

// Auxiliary per-handle device arrays read by the stencils.
struct StencilAux {
    const uint8_t* ct;      // cell types (has_solid)
    const float* fx;        // face fractions, padded layout (has_closed)
    const float* fy;
    const float* fz;
    const float* bu;        // per-cell buoyancy 1/rho (hetero)
    const float* kap;       // per-cell K*dt/dx (hetero)
    const float* xi;        // per-cell wall term sum (hetero + impedance; edges only)
    const float* alpha;     // alpha(d), d = 0..n_pml
    const float* pdamp;     // 1/(1+alpha(d) dt)
    const int* status;      // nonzero: a previous step failed, skip work
};

struct AcousticFDTDState {
    int Nx, Ny, Nz, N, Nplate, plate_iz;
    int Nvx, Nvy, Nvz;
    int order, boundary, n_pml, storage, family;
    int hetero, has_solid, has_frac;
    float dx, dt, c, rho_air;
    StencilParams sp;
    hipStream_t stream;
    hipEvent_t ev0, ev1;

    // Fields (element type per `storage`; bytes per element in esz; bytes per
    // field buffer in fbytes, which for block floating point includes the
    // row exponents after the mantissas).
    int esz;
    size_t fbytes;
    void* P[2];             // P[cur] = P^n
    int cur;
    void* V[3];             // V^{n-1/2} (current)
    void* V2[3];            // second velocity buffer of the fused sweep
    // Conversion staging for non-FP32 storage.
    float* stage;

    // Stencil aux arrays.
    uint8_t* d_ct;
    float *d_fx, *d_fy, *d_fz;
    float *d_bu, *d_kap, *d_xi;
    float *d_alpha, *d_pdamp;
    int* d_status;          // [0] code, [1] step index at failure
    std::vector<uint8_t> h_ct;   // cell types after retyping (host copy)

    // Kirchhoff plate.
    int nact;
    float *d_w[3];          // w_curr, w_prev, w_tmp rotate
    float* d_ext;
    int *d_act, *d_above, *d_below, *d_b13, *d_vzb, *d_vza;
    float plate_a, plate_bh, plate_vd, plate_bk;
    std::vector<uint8_t> h_pa;

    // Bridge sources.
    int nbsrc;
    int* d_bidx;
    float* d_bdrive;        // staging for the bridge drives (nbsrc)
    int* d_bgrp;            // bridge entries grouped by plate node: start | node | entry
    int nbgrp;
    std::vector<int> h_bidx;
    std::vector<float> h_bwgt;
    int* d_fidx;            // plate-force staging: grouped entries (start | node | entry)
    float* d_fval;          // plate-force staging: weights
    int fcap;

    // Point sources.
    int npsrc, plen;
    int* d_pidx;
    float* d_psig;

    // Receivers.
    int nrec;
    int* d_ridx;
    float* d_rwgt;
    float* d_trace;
    int trace_cap, trace_n;

    // Scratch for API-side gathers.
    int* d_gidx;
    float* d_gwgt;
    float* d_gout;
    int gcap;
    // This device's copy of the file-scope scratch array of the
    // allocation-free sampler (resolved at creation, never freed).
    unsigned char* d_scratch;

    int step_count, phase;
    // Hooked steps (afdtd_step_hooked): the hook's kernel clears the plate
    // force accumulator, so every step but the call's last skips its fill.
    int hook_clears_ext;
    int fused_tile_y, fused_seg_x;
    // Single-step sweep with cell geometry: per-cell code bytes (walls,
    // closed / fractional / plate faces), rebuilt when the fractions change.
    int fused_geo;
    uint8_t* d_geo;
    float* d_vpl;           // plate velocity of the current step (nact > 0)
    // Two-step sweep with cell geometry (order 4, FP32, homogeneous): per-cell
    // pair codes, step n+1's plate terms [5][Nplate] and P^{n+1} beside the
    // plate [Nplate][2].
    int geo_pair;
    int kernel_opt;         // the kernel family the handle was created with
    uint8_t* d_g2;
    uint32_t* d_g2p;        // the pair codes, four planes per word
    float* d_tw;            // [3][N] divergence-tap weights (fractions present)
    int* d_g2_items;        // work items: nfast without weighted taps, then nslow with
    int g2_nfast, g2_nslow, g2_items_seg, g2_gen, g2_items_gen;
    hipStream_t g2_stream;  // the weighted-tap tiles' stream (default priority)
    hipEvent_t g2_ev[2];
    float* d_ppre;
    float* d_pab;
    // Side job of the geometry sweep (set by a hooked caller, the co-evolver):
    // one extra workgroup forms sum_i w[idx[i]] * wgt[i] over the plate
    // displacement the sweep's step ends with, in index order from +0, and
    // stores it with the step index as tag.
    const int* side_idx;
    const float* side_wgt;
    int side_n;
    float* side_out;        // device: [0] the sum, [1] the step index (int bits)
    int side_step;          // step index of the sweep being enqueued

    // Two-step sweep: host copies of source cells and receiver taps, and the
    // per-tile tables built from them (rebuilt when f2_dirty).
    std::vector<int> h_pidx, h_ridx;
    int* d_f2_src_start;
    void* d_f2_src;
    int* d_f2_tap_start;
    void* d_f2_tap;
    float* d_f2_tapbuf;
    size_t f2_tapcap;
    int f2_dirty;
    int f2_ok;              // tables fit the per-tile caches (else single steps)
    int f2_rows;            // extended rows per thread: 2 (16-row tiles) or 1 (8-row)
};

// Steps with a per-step hook: hook(q, s) runs before step q (global index s)
// is enqueued and once more with q == n_steps after the last one, and
// enqueues its own work on st->stream; every step is a single-step launch.
// Step bookkeeping and failure handling are those of fdtd_step.  A hooked
// kernel that finds its own failure sets d_status = {FDTD_STATUS_HOOK, s}
// (s = the step that must not run); status_out receives the raw status word
// read at the end of the call ({0, 0} when nothing failed).
#include <functional>
using StepHook = std::function<int(int q, int s)>;
constexpr int FDTD_STATUS_HOOK = 16;
int afdtd_step_hooked(AcousticFDTDState* st, int n_steps, const StepHook& hook, float* ms_total,
                      int* status_out);

// The same for the AMR stepper (afdtd_amr.h).  hook(a) runs before step
// a.q of the call (a.q < n) and once more with a.q == n; it sees the plate
// displacement current at that point (the buffer rotation advances by one
// per step) and the device buffers the co-evolver's kernel reads and writes.
// A hooked kernel that finds its own failure before step q sets status[3]
// = q and lowers status[0] to q - 1, so that step q and everything after it
// is skipped.  Returns SK_OK, SK_ERR_DIVERGED (a step diverged, or the hook
// failed: *hook_failed) or AMR_ERR_DEVICE; *done = the steps that completed.
#include "../../include/afdtd_amr.h"
#include "../../include/afdtd_coevolver.h"
struct AmrHookArgs {
    int q;
    const float* w;         // current plate displacement
    float* ext;             // plate force accumulator
    const float* p;         // cell pressures
    const float* v;         // face velocities
};
using AmrStepHook = std::function<int(const AmrHookArgs&)>;
int amr_step_hooked(AcousticAMRState* st, int n, const AmrStepHook& hook, float* ms_total,
                    int* done, int* hook_failed);
// What the co-evolver's kernel needs of an AMR state: its stream, status
// words, plate size and the device mic stencils (CSR: per-mic starts, then
// cells / faces and weights).
struct AmrCoView {
    hipStream_t stream;
    int* status;
    int n_cells, n_pnodes, pNx, pNy, n_mics;
    int n_mp, n_mv;                         // stencil entries: pressure cells, velocity faces
    const int* mp_start;
    const int* mp_cell;
    const float* mp_w;
    const int* mv_start;
    const int* mv_face;
    const float* mv_w[3];
};
int amr_co_view(AcousticAMRState* st, AmrCoView* v);
// The AMR state of amr_pressure_backend_create (the same checks, calls,
// error codes and messages), without the interface object around it.
AcousticAMRState* amr_state_from_descriptor(const AMRCoevolverDescriptor* d);

// Launchers (all enqueue on st->stream; return hipError_t of the launch).
hipError_t launch_velocity_generic(AcousticFDTDState* st);
hipError_t launch_pressure_generic(AcousticFDTDState* st);
hipError_t launch_fused_step(AcousticFDTDState* st);
bool fused_supported(const AcousticFDTDState* st);
// The single-step sweep with cell geometry (FP32 storage).
bool fused_geo_supported(const AcousticFDTDState* st);
// Build st->d_geo from the cell types, face fractions and plate faces.
hipError_t launch_geo_codes(AcousticFDTDState* st);
// Two steps (global indices step, step+1) in one sweep; the receiver taps of
// the first step go to a tap buffer that launch_fused2_taps turns into a
// trace row.
hipError_t launch_fused2_pair(AcousticFDTDState* st, int step);
hipError_t launch_fused2_taps(AcousticFDTDState* st, int trace_row);
// (Re)build the per-tile source / tap tables when sources or receivers
// changed; returns FDTD_OK and sets st->f2_ok.
int fused2_prepare(AcousticFDTDState* st);
bool fused2_supported(const AcousticFDTDState* st);
// Trace row of `nrec` 8-tap receivers from a tap buffer (w0*p0 + ... + w7*p7).
hipError_t launch_tap_trace(const float* tapbuf, const float* rwgt, int nrec, float* trace,
                            int trace_row, const int* status, hipStream_t s);
// Entries (key, id) of the per-tile source and tap lists of one grid for the
// two-step sweep with TYO-row tiles: list t = ty + nty*tz, unsorted.
void f2_tile_lists(int Nx, int Ny, int Nz, int TYO, int nsrc, const int* src_cell,
                   const int* src_id, int ntap, const int* tap_cell, const int* tap_slot,
                   std::vector<std::vector<int2>>& src, std::vector<std::vector<int2>>& tap);
// Own rows of a two-step tile for tile_y 8 or 16.
int f2_rows_for(int tile_y);   // two-step sweep: rows per thread for a tile height
// Two-step sweep of the order-4 stencil (afdtd_fused4.hip): FP32, homogeneous
// medium, rigid walls / PML, no walls, fractions or plate.
bool fused4_supported(const AcousticFDTDState* st);
int f4_tile_rows();
void f4_tile_lists(int Ny, int Nz, int nsrc, const int* src_cell, const int* src_id, int ntap,
                   const int* tap_cell, const int* tap_slot, std::vector<std::vector<int2>>& src,
                   std::vector<std::vector<int2>>& tap);
hipError_t launch_fused4_pair(AcousticFDTDState* st, int step);
// The two-step sweep with cell geometry (walls, closed / fractional faces,
// active plate): order 4, FP32, homogeneous medium.
bool fused4_geo_supported(const AcousticFDTDState* st);
// Build st->d_g2 from st->d_geo and the plate (after launch_geo_codes).
hipError_t launch_geo_pair_codes(AcousticFDTDState* st);
// Steps s, s+1 with cell geometry: the step-s plate kernel (and its force
// fill) has run and the plate buffers are rotated; this enqueues step s+1's
// plate terms, the sweep, the step-s receiver taps (trace_row0), step s+1's
// plate update (force fill and rotation included) and, if that update trips
// the plate guard, step s again as one single step over the sweep's output,
// leaving status {2, s+1}.
hipError_t launch_fused4_geo_pair(AcousticFDTDState* st, int step, int trace_row0);
// Single-step sweep that runs only while the status word equals run_on.
hipError_t launch_fused_step_if(AcousticFDTDState* st, int run_on);
int f2_tile_rows(int tile_y);
// Longest per-tile source / tap list the two-step sweep caches in LDS.
constexpr int F2_MAX_SRC = 32, F2_MAX_TAP = 256;

// One room of a batch (afdtd_batch.hip): its grid, its work items
// [tile_base, tile_base + nty*ntz*nseg) and, for the two-step sweep, its
// per-tile source / tap lists [list_base, list_base + nty*ntz).
struct RoomDesc {
    int Nx, Ny, Nz, nty, ntz, nseg, tile_base, list_base;
    long long off;          // element offset of the room in every field array
    float dt, vel_coeff, bulk_coeff;
    float gside[6];
    float xterm[6];         // wall terms (kap/2) * g / sqrt(rho K)
};
// Two steps of every room of a batch in one launch of the two-step sweep.
struct F2BatchLaunch {
    StencilParams common;   // shared parameters (per-room fields from RoomDesc)
    const void* Pin;        // room-0 bases of the field arrays
    void* Pout;
    const void* Vin[3];
    void* Vout[3];
    const RoomDesc* rooms;
    const int* tile_room;   // room of every work item
    int ntiles, seg, tile_y, storage;
    const float* alpha;
    const float* pdamp;
    const int* status;
    const int* src_start;   // per list: entries [src_start[l], src_start[l+1])
    const int2* src;        // (key, room), sorted by key
    const int* tap_start;
    const int2* tap;        // (key, tap slot), sorted by key
    const float* psig;      // rooms x plen source samples
    int plen, step;
    float* tapbuf;
    hipStream_t stream;
};
hipError_t launch_fused2_batch(const F2BatchLaunch& L);
hipError_t launch_plate_step(AcousticFDTDState* st, int step_index);
hipError_t launch_plate_inject(AcousticFDTDState* st);
hipError_t launch_post_step(AcousticFDTDState* st, int step_index, int trace_row);
hipError_t launch_stability_check(AcousticFDTDState* st, int start, int step_index);
hipError_t launch_fill(float* p, size_t n, float v, hipStream_t s);
// Force injection grouped by target node (grp = start[ng+1] | node[ng] |
// entry[...]): one thread per node adds its entries in call order.
hipError_t launch_accumulate_groups(float* ext, const int* grp, int ng, const float* val,
                                    float scale, float inv_dx2, const int* status,
                                    hipStream_t s);
// Ncells / Nz / n locate the per-row exponents of block-floating-point fields
// (ignored for the other storage types).
hipError_t launch_gather(const void* F, int storage, float scale, const int* idx,
                         const float* wgt, int n, int taps, int zero_start, float* out,
                         hipStream_t s, int Ncells, int Nz);
hipError_t launch_to_f32(const void* src, int storage, float scale, float* dst, size_t n,
                         hipStream_t s, int Nz);
hipError_t launch_from_f32(const float* src, void* dst, int storage, float scale, size_t n,
                           hipStream_t s, int Nz);
hipError_t launch_energy(AcousticFDTDState* st, double* d_out);
