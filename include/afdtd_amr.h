/*
 * afdtd_amr.h — C ABI of the MI355X AMR (topology-driven) acoustic stepper
 * in libafdtd.so.
 *
 * Part 1 declares the reference's AMR functions (csrc/include/acoustic_amr.h)
 * that cover the gradient-order-2 stepper, with the same names, signatures,
 * return codes and ownership rules; each cites the declaration it replaces
 * (csrc/include/acoustic_amr.h:line) and the definition whose behaviour it
 * reproduces (csrc/kernels/acoustic_amr.cpp:line).  Both gradient orders
 * are provided: 2 (face-pair difference) and 8 (per-face 8-cell stencil
 * built at create with the reference's weight recurrence, as written).
 *
 * Part 2 declares extensions under new names (state upload, timed steps).
 *
 * Topology: n_cells pressure cells and n_faces velocity faces; face f joins
 * face_cell_neg[f] -> face_cell_pos[f], with area, open fraction and
 * centre distance.  Velocity is positive from neg to pos.
 *
 * Device execution: every handle owns one HIP stream.  amr_step enqueues two
 * kernels per step (faces, then cells; with a plate also the plate's face
 * condition and the plate step) and synchronises once before returning (it
 * must report SK_ERR_DIVERGED).  Getters are synchronous on
 * return.  Handles are not thread-safe (as in the reference).
 */
#ifndef AFDTD_AMR_H
#define AFDTD_AMR_H

#include <stdint.h>

#if defined(_WIN32)
#  define SK_API
#else
#  define SK_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes — reference csrc/include/serial_kernel.h:47-50 */
#ifndef SK_OK
#define SK_OK                0
#define SK_ERR_NULL_STATE   (-1)
#define SK_ERR_DIM_MISMATCH (-2)
#define SK_ERR_DIVERGED     (-3)
#endif
/* Extension code: HIP runtime error (never returned by a reference path). */
#define AMR_ERR_DEVICE      (-4)

/* Border modes — reference csrc/include/acoustic_amr.h:111-116 */
#define AMR_BORDER_ANECHOIC      0
#define AMR_BORDER_ROOM_PANEL    1
#define AMR_BORDER_TRANSMISSION  2

typedef struct AcousticAMRState AcousticAMRState;

/* ═════════════════════ Part 1: the reference ABI ═══════════════════════ */

/* acoustic_amr.h:20-35; acoustic_amr.cpp:682-880 (inputs validated as
 * :275-311, copied; dt = 0.77*min_dx/(c*sqrt 3); per-face 1/d and area*open,
 * per-cell 1/(V*max(open,1e-12)) on acoustic cells, signed CSR of faces per
 * cell in face order; gradient order 8 also builds each face's stencil:
 * up to 4 cells on each side along the face normal, walked through
 * axis-neighbour tables or the most aligned face, weighted by the
 * reference's recurrence (acoustic_amr.cpp:313-679), else the two-point
 * difference; dt uses 0.10 instead of 0.77 for order 8). */
SK_API AcousticAMRState* amr_create(
    int            n_cells,
    const double*  cell_centers,
    const double*  cell_volumes,
    const double*  open_volume_frac,
    const uint8_t* cell_types,
    int            n_faces,
    const int32_t* face_cell_neg,
    const int32_t* face_cell_pos,
    const double*  face_area,
    const double*  face_open_frac,
    const double*  face_distance,
    double         c,
    double         rho_air,
    double         min_dx,
    int            gradient_order);

/* acoustic_amr.h:44-45; acoustic_amr.cpp:897-900, :917-920 */
SK_API const char* amr_get_last_create_error_stage(void);
SK_API const char* amr_get_last_create_error_message(void);

/* acoustic_amr.h:53-55; acoustic_amr.cpp:902-915.  The order-8 stencil is
 * built synchronously inside amr_create, so "active" reads 0; done and total
 * faces report the last order-8 build. */
SK_API int amr_get_create_progress_active(void);
SK_API int amr_get_create_progress_done_faces(void);
SK_API int amr_get_create_progress_total_faces(void);

/* acoustic_amr.h:57; acoustic_amr.cpp:922 (NULL-safe) */
SK_API void amr_destroy(AcousticAMRState* st);

/* acoustic_amr.h:59; acoustic_amr.cpp:1100-1225.  Per step: face velocity
 * update, fused divergence + pressure update, wall/plate cells zeroed,
 * non-finite pressure or (every 256th step) |p| > 2e3 -> SK_ERR_DIVERGED
 * with step_count left at the failing step. */
SK_API int amr_step(AcousticAMRState* st, int n_steps);

/* acoustic_amr.h:60; acoustic_amr.cpp:924-936 */
SK_API int amr_reset(AcousticAMRState* st);

/* acoustic_amr.h:62-65; acoustic_amr.cpp:1227-1244 (nearest acoustic cell
 * by squared distance, first in index order on ties; non-finite value ->
 * SK_ERR_DIVERGED) */
SK_API int amr_inject_pressure_nearest(AcousticAMRState* st, const double* xyz, float value);

/* acoustic_amr.h:67-70; acoustic_amr.cpp:1246-1252 (out_count == n_cells) */
SK_API int amr_get_pressure(const AcousticAMRState* st, float* out_pressure, int out_count);

/* acoustic_amr.h:78-81; acoustic_amr.cpp:1254-1264 (n_cells_3 >= 3*n_cells) */
SK_API int amr_get_cell_centers(const AcousticAMRState* st, float* out_xyz, int n_cells_3);

/* acoustic_amr.h:83-86; acoustic_amr.cpp:1296-1302 (out_count == n_faces) */
SK_API int amr_get_velocity(const AcousticAMRState* st, float* out_velocity, int out_count);

/* acoustic_amr.h:88-89; acoustic_amr.cpp:1304-1312 */
SK_API double amr_get_dt(const AcousticAMRState* st);
SK_API int amr_get_step_count(const AcousticAMRState* st);

/* acoustic_amr.h:102-106; acoustic_amr.cpp:1266-1294 (nearest voxel by
 * lround, last cell in index order wins, zero elsewhere) */
SK_API int amr_scatter_pressure_uniform(const AcousticAMRState* st, int Nx, int Ny, int Nz,
                                        const double bmin[3], const double bmax[3], float* out);

/* acoustic_amr.h:132-140; acoustic_amr.cpp:1318-1409 (graded sigma on
 * PML-type cells by distance to the bounding box, P_damp = exp(-s dt),
 * P_src = (1-exp(-s dt))/(s dt), face damping from the open-volume-weighted
 * harmonic mean; n_pml <= 0 leaves the identity) */
SK_API int amr_set_border_condition(AcousticAMRState* st, int mode, float sigma_order,
                                    float R_reflection, float border_Z_match, int n_pml,
                                    const double* bounds_min, const double* bounds_max);

/* acoustic_amr.h:165-186; acoustic_amr.cpp:1412-1531.  Kirchhoff plate on a
 * plate_Nx x plate_Ny grid (flat index j + plate_Ny*i) coupled to AMR faces
 * above (neg = plate side, air on pos) and below (air on neg), or to one
 * fallback cell per side.  Each step: the coupled faces take the plate
 * velocity (+-(w - w_prev)/dt) after the face update; after the pressure
 * update the plate advances (13-point biharmonic, Rayleigh damping, acoustic
 * load (P_below - P_above)*dx^2, external force, which is then cleared).
 * Besides the reference's checks, active flat indices outside the plate
 * grid are rejected (SK_ERR_DIM_MISMATCH); they must be distinct. */
SK_API int amr_setup_plate(AcousticAMRState* st, int plate_Nx, int plate_Ny, float plate_dx,
                           const float plate_origin[3], const uint8_t* plate_active,
                           float plate_rho_h, float plate_D, float plate_alpha_M,
                           float plate_beta_K, int n_active_plate,
                           const int32_t* plate_active_flat_idx,
                           const int32_t* face_above_starts, const int32_t* face_above_idx,
                           const float* face_above_wgt, const int32_t* face_below_starts,
                           const int32_t* face_below_idx, const float* face_below_wgt,
                           const int32_t* cell_above, const int32_t* cell_below);

/* acoustic_amr.h:195-210; acoustic_amr.cpp:1534-1563 */
SK_API int amr_set_bridge_plate_sources(AcousticAMRState* st, int n, const int32_t* idx,
                                        const float* wgt);
SK_API int amr_set_neck_plate_sources(AcousticAMRState* st, int n, const int32_t* idx,
                                      const float* wgt);

/* acoustic_amr.h:226-244; acoustic_amr.cpp:1567-1597 (external force
 * accumulated in call order; an out-of-range node stops the loop with
 * SK_ERR_DIM_MISMATCH after the entries before it were applied) */
SK_API int amr_inject_bridge_drive(AcousticAMRState* st, const float* cell_drives, int n_cells,
                                   float force_scale);
SK_API int amr_inject_plate_force(AcousticAMRState* st, int n, const int32_t* plate_idx,
                                  const float* wgt, float total_force_N);

/* acoustic_amr.h:256-283; acoustic_amr.cpp:1600-1644 */
SK_API int amr_sample_plate_displacement_batch(const AcousticAMRState* st, int n_points,
                                               const int* idx4, const float* wgt4,
                                               float* out_w);
SK_API int amr_sample_plate_weighted(const AcousticAMRState* st, int n, const int* plate_idx,
                                     const float* wgt, float* out_w);
SK_API int amr_get_plate_displacement(const AcousticAMRState* st, float* out, int out_len);

/* acoustic_amr.h:301-325; acoustic_amr.cpp:1647-1792.  Pressure: inverse-
 * distance weights over acoustic cells within 2*cbrt(mean acoustic volume)
 * of the (envelope-clamped) position; velocity (need_velocity != 0):
 * area-weighted face normals within the same radius.  A second set-up call
 * appends to each mic's lists, as the reference does. */
SK_API int amr_setup_mic_samplers(AcousticAMRState* st, int n_mics, const float* pos_xyz,
                                  const float* axis_xyz, int need_velocity);
SK_API int amr_sample_mics(const AcousticAMRState* st, int n_mics, float* out_p, float* out_vx,
                           float* out_vy, float* out_vz);

/* ═════════════════════ Part 2: extensions (no reference counterpart) ═══ */

/* Upload a state: p (n_cells floats) and/or v (n_faces floats); NULL keeps
 * that field.  For tests and warm starts. */
SK_API int amr_set_state(AcousticAMRState* st, const float* p, const float* v);

/* amr_step with HIP-event timing on the handle's stream: *ms_total = all
 * n_steps, *ms_faces / *ms_cells = summed duration of the face and cell
 * kernels (events around each launch). */
SK_API int amr_step_timed(AcousticAMRState* st, int n_steps, float* ms_total,
                          float* ms_faces, float* ms_cells);

/* Number of CSR (cell, face) entries = 2 * n_faces; for byte accounting. */
SK_API int amr_get_csr_entries(const AcousticAMRState* st);

/* Record layout of the step kernels.  Direct: every face and cell streams its
 * full record (16 B per face and cell; order 8 adds 64 B of stencil per
 * face).  Compact: each face streams {neg cell, class} (8 B) and each cell a
 * class id (4 B) into tables of the distinct records, bit for bit the same
 * values.  layout 0 = automatic (compact while each table has at most 65536
 * records), 1 = direct, 2 = compact up to 2^24 records.  Rebuilt on every
 * coefficient change (border condition).  amr_get_record_layout returns the
 * layout in use (1 direct, 2 compact). */
SK_API int amr_set_record_layout(AcousticAMRState* st, int layout);
SK_API int amr_get_record_layout(const AcousticAMRState* st);

/* The acoustic half of the co-evolver's per-sample loop on the device:
 * n_samples audio samples of `substeps` steps each; after each sample every
 * mic (amr_setup_mic_samplers) records (p*a + b*rho_c*(v . axis)) * gain,
 * the co-evolver's polar mix (reference acoustic_coevolver.cpp:2208-2219),
 * into out[sample * n_mics + mic].  One host synchronisation per call.
 * polar_a, polar_b, gain: n_mics floats; axis_xyz: 3 per mic.  On a
 * divergence returns SK_ERR_DIVERGED with *n_done = samples completed
 * before the failing step (later rows are 0). */
SK_API int amr_record_mic_samples(AcousticAMRState* st, int n_samples, int substeps,
                                  const float* polar_a, const float* polar_b,
                                  const float* axis_xyz, const float* gain, float rho_c,
                                  float* out, int* n_done);

#ifdef __cplusplus
}
#endif

#endif /* AFDTD_AMR_H */
