/*
 * afdtd_coevolver.h — C ABI of the instrument co-evolver in libafdtd.so:
 * strings, bridge and neck impedances, pickups and microphones coupled to
 * a body (the uniform-grid FDTD of afdtd.h, or the AMR stepper of
 * afdtd_amr.h through the pressure-backend descriptor), advanced per audio
 * sample.
 *
 * Each declaration has the name, signature, return codes and ownership of
 * the reference's co-evolver ABI (csrc/include/acoustic_coevolver.h:line)
 * and reproduces the behaviour of its definition
 * (csrc/kernels/acoustic_coevolver.cpp:line).  Descriptor structs keep the
 * reference's field order and types, so a caller that fills them for the
 * reference fills them for this library.
 *
 * Execution: the whole per-sample loop runs on the device.  One
 * coevolver_step call enqueues, per body sub-step, one single-workgroup
 * kernel (neck mode, every string's sub-steps, bridge impedances, plate
 * loads; the previous sub-step's saddle gather, modal projection and, at a
 * sample's end, pickups and microphones) followed by the body FDTD step,
 * and synchronises once at the end.  No host round trip happens between
 * sub-steps or samples.  Outputs land in device ring buffers of 4096
 * samples per pickup and microphone.
 *
 * Differences from the reference, all at its undefined or unreachable edges:
 *   - at most 64 strings (the reference indexes fixed 64-entry arrays);
 *   - at most 4096 string nodes in total (the strings live in one
 *     workgroup's LDS during a sub-step);
 *   - after CE_ERR_UNSTABLE the failing sample's partial string state is not
 *     reproduced (the sample counter, output rings and body fields are).
 * Handles are not thread-safe, as in the reference; the async calls run the
 * same device loop from one worker thread.
 */
#ifndef AFDTD_COEVOLVER_H
#define AFDTD_COEVOLVER_H

#include <stdint.h>

#include "afdtd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes — reference csrc/include/acoustic_coevolver.h:115-129 */
#define CE_OK               0
#define CE_ERR_NULL        (-1)
#define CE_ERR_DIM         (-2)
#define CE_ERR_UNSTABLE    (-3)
#define CE_ERR_PARAM       (-4)
#define CE_ERR_STRING_IDX  (-5)
#define CE_ERR_BUSY        (-6)
#define CE_ERR_FDTD        (-7)
#define CE_ERR_FDTD_BRIDGE (-8)
#define CE_ERR_FDTD_STEP   (-9)
#define CE_ERR_FDTD_MIC_P  (-10)
#define CE_ERR_FDTD_MIC_V  (-11)

/* Async worker status — reference acoustic_coevolver.h:133-137 */
#define CE_STATUS_IDLE      0
#define CE_STATUS_RUNNING   1
#define CE_STATUS_DONE      2
#define CE_STATUS_CANCELLED 3
#define CE_STATUS_ERROR     4

/* Progress snapshot — reference acoustic_coevolver.h:146-151 */
typedef struct {
    int samples_done;
    int samples_total;
    int status;
    int error_code;
} CoEvolverProgress;

/* Pickup types — reference acoustic_coevolver.h:156-160 */
#define PICKUP_SINGLE_COIL 0
#define PICKUP_HUMBUCKER   1
#define PICKUP_PIEZO       2

/* String descriptor — reference acoustic_coevolver.h:175-199.
 * path_xyz: (n_segs+1, 3) world-space nodes, nut (0) to saddle (n_segs). */
typedef struct {
    int          n_segs;
    const float* path_xyz;
    float        tension_N;
    float        linear_mass_kgm;
    float        damping;
    float        stiffness_EI;
    float        axial_stiffness_N;
    float        neck_freq_hz;
    float        neck_mass_kg;
    float        neck_Q;
} CoEvolverStringDef;

/* Pickup descriptor — reference acoustic_coevolver.h:217-225 */
typedef struct {
    int      type;
    float    pos[3];
    float    axis[3];
    float    pole_sigma;
    float    coil_spacing;
    float    sensitivity;
    uint32_t string_mask;
} CoEvolverPickupDef;

/* Microphone descriptor — reference acoustic_coevolver.h:246-252:
 * output = gain * (polar_a * P + polar_b * rho_air * c * dot(v, axis)). */
typedef struct {
    float pos[3];
    float axis[3];
    float polar_a;
    float polar_b;
    float gain;
} CoEvolverMicDef;

/* Body descriptor — reference acoustic_coevolver.h:264-282 */
typedef struct {
    int            Nx, Ny, Nz;
    float          dx;
    float          origin[3];
    float          c;
    float          rho_air;
    const uint8_t* cell_type;
    int            plate_iz;
    const uint8_t* plate_active;
    float          plate_mass_density;
    float          plate_stiffness_D;
    float          plate_alpha_M;
    float          plate_beta_K;
    int            n_pml;
    int            n_bridge_src;
    const float*   bridge_src_xyz;
    float          bridge_sigma;
} CoEvolverBodyDef;

#ifndef AFDTD_HAVE_AMR_COEVOLVER_DESCRIPTOR
#define AFDTD_HAVE_AMR_COEVOLVER_DESCRIPTOR
/* AMR body descriptor — reference acoustic_coevolver.h:325-397: field order
 * and types of the reference's struct.  The arrays are copied by the
 * constructor and need only outlive the call. */
typedef struct AMRCoevolverDescriptor {
    int            n_cells;
    const double*  cell_centers;
    const double*  cell_volumes;
    const double*  open_volume_frac;
    const uint8_t* cell_types;
    const int16_t* cell_levels;

    int            n_faces;
    const int32_t* face_cell_neg;
    const int32_t* face_cell_pos;
    const double*  face_area;
    const double*  face_open_frac;
    const double*  face_distance;

    double c;
    double rho_air;
    double min_dx;

    double bounds_min[3];
    double bounds_max[3];

    int            plate_Nx;
    int            plate_Ny;
    float          plate_dx;
    float          plate_origin[3];
    const uint8_t* plate_active;
    float          plate_mass_density;
    float          plate_stiffness_D;
    float          plate_alpha_M;
    float          plate_beta_K;

    int            n_active_plate;
    const int32_t* plate_active_flat_idx;
    const int32_t* plate_face_above_starts;
    const int32_t* plate_face_above_idx;
    const float*   plate_face_above_wgt;
    const int32_t* plate_face_below_starts;
    const int32_t* plate_face_below_idx;
    const float*   plate_face_below_wgt;
    const int32_t* plate_cell_above;
    const int32_t* plate_cell_below;

    int            n_bridge_plate;
    const int32_t* bridge_plate_idx;
    const float*   bridge_plate_wgt;

    int            n_neck_plate;
    const int32_t* neck_plate_idx;
    const float*   neck_plate_wgt;

    double soundhole_cx;
    double soundhole_cy;
    double soundhole_radius;

    int   border_mode;
    float border_sigma_order;
    float border_R_reflection;
    float border_Z_match;
    int   n_pml;
    int   gradient_order;
} AMRCoevolverDescriptor;
#endif

typedef struct AcousticCoEvolverState AcousticCoEvolverState;

/* Construction — reference acoustic_coevolver.h:306-318, definition
 * acoustic_coevolver.cpp:987-1381 (body FDTD with the canonical soundhole
 * carved from the plate mask, bridge kernel, strings with per-string
 * sub-stepping, shared neck mode, pickup B-kernels, mic samplers, saddle
 * bilinear taps, per-string bridge kernels).  NULL on bad arguments. */
SK_API AcousticCoEvolverState* coevolver_create(
    int n_strings, const CoEvolverStringDef* string_defs,
    int n_pickups, const CoEvolverPickupDef* pickup_defs,
    int n_mics, const CoEvolverMicDef* mic_defs,
    const CoEvolverBodyDef* body, float sample_rate, int modal_stride, float force_scale);

/* AMR-backed construction — reference acoustic_coevolver.h:419-429,
 * definition acoustic_coevolver.cpp:1387-1746: the body is the AMR
 * pressure backend built from the descriptor (amr_pressure_backend_create:
 * CSR stepper, plate, bridge and optional neck plate sources, border
 * condition); the plate grid stands in for the uniform grid (saddle taps,
 * bridge kernels with sigma 2 * plate_dx over the bridge plate nodes); mics
 * sample through the backend's stencils; the neck mode is off without neck
 * plate nodes.  NULL on failure, with the stage label in
 * coevolver_create_amr_last_error(). */
SK_API AcousticCoEvolverState* coevolver_create_amr(
    int n_strings, const CoEvolverStringDef* string_defs,
    int n_pickups, const CoEvolverPickupDef* pickup_defs,
    int n_mics, const CoEvolverMicDef* mic_defs,
    const AMRCoevolverDescriptor* amr_desc,
    float sample_rate, int modal_stride);

/* Stage label of the last coevolver_create_amr failure on this thread, ""
 * after a success — reference acoustic_coevolver.h:431-432,
 * acoustic_coevolver.cpp:1749-1752. */
SK_API const char* coevolver_create_amr_last_error(void);

/* Extensions (no reference counterpart as C symbols): the AMR pressure
 * backend's error code name and message of the last failed
 * coevolver_create_amr on this thread, for callers without C++ linkage. */
SK_API const char* coevolver_amr_backend_error_code(void);
SK_API const char* coevolver_amr_backend_error_message(void);

/* reference acoustic_coevolver.h:320, acoustic_coevolver.cpp:1797-1890 */
SK_API void coevolver_destroy(AcousticCoEvolverState* st);

/* Excitation — reference acoustic_coevolver.h:455-489,
 * acoustic_coevolver.cpp:2413-2509 (raised-cosine draw ramp; pre-scaled
 * per-sample force buffer at the nearest interior node). */
SK_API int coevolver_pluck_string(AcousticCoEvolverState* st, int string_idx,
                                  float position_norm, float amplitude, int n_duration_samples);
SK_API int coevolver_inject_string_force(AcousticCoEvolverState* st, int string_idx,
                                         int n_samples, const float* force_buf,
                                         float force_pos_norm);

/* Pluck schedule — reference acoustic_coevolver.h:507-521,
 * acoustic_coevolver.cpp:1756-1795 */
SK_API int coevolver_schedule_pluck(AcousticCoEvolverState* st, int onset_sample, int string_idx,
                                    float position_norm, float amplitude, int n_duration_samples);
SK_API void coevolver_clear_pluck_schedule(AcousticCoEvolverState* st);

/* Main step — reference acoustic_coevolver.h:547, acoustic_coevolver.cpp:1910-2251 */
SK_API int coevolver_step(AcousticCoEvolverState* st, int n_samples);

/* reference acoustic_coevolver.h:574-581, acoustic_coevolver.cpp:2253-2321 */
SK_API int coevolver_step_block_with_drive(AcousticCoEvolverState* st,
                                           const float* const* string_drive_blocks,
                                           int n_strings, int n_samples, float force_pos_norm,
                                           float* mic_out_buf);

/* Async stepping — reference acoustic_coevolver.h:592-616,
 * acoustic_coevolver.cpp:2323-2411 */
SK_API int coevolver_step_async(AcousticCoEvolverState* st, int n_samples);
SK_API void coevolver_get_progress(const AcousticCoEvolverState* st, CoEvolverProgress* out);
SK_API int coevolver_wait(AcousticCoEvolverState* st);
SK_API int coevolver_cancel(AcousticCoEvolverState* st);
SK_API int coevolver_is_running(const AcousticCoEvolverState* st);

/* Outputs — reference acoustic_coevolver.h:628-818, acoustic_coevolver.cpp:2511-2696.
 * On an AMR body, as there: the pressure field is the flat cell array, the
 * uniform variant scatters it onto a box grid, the plate getters read the
 * plate grid, coevolver_get_amr_cell_centers returns the cell centres
 * (CE_ERR_PARAM on a uniform body), coevolver_sample_pressure returns
 * CE_ERR_PARAM, and the calls that need a uniform grid
 * (coevolver_set_face_fractions, coevolver_get_surface_emission,
 * coevolver_sample_velocity) return CE_ERR_NULL. */
SK_API int coevolver_get_pickup_output(const AcousticCoEvolverState* st, int pickup_idx,
                                       float* out, int n_samples);
SK_API int coevolver_get_mic_output(const AcousticCoEvolverState* st, int mic_idx, float* out,
                                    int n_samples);
SK_API int coevolver_get_pressure_field(const AcousticCoEvolverState* st, float* out,
                                        int out_len);
SK_API int coevolver_get_pressure_field_uniform(const AcousticCoEvolverState* st, int Nx,
                                                int Ny, int Nz, const double bmin[3],
                                                const double bmax[3], float* out, int out_len);
SK_API int coevolver_get_amr_cell_centers(const AcousticCoEvolverState* st, float* out_xyz,
                                          int n_cells_3);
SK_API int coevolver_get_plate_displacement(const AcousticCoEvolverState* st, float* out,
                                            int out_len);
SK_API int coevolver_get_plate_dims(const AcousticCoEvolverState* st, int* out_Nx, int* out_Ny,
                                    int* out_count);
SK_API int coevolver_get_modal_amplitudes(const AcousticCoEvolverState* st, int string_idx,
                                          float* out_re, float* out_im, int n_modes);
SK_API int coevolver_sample_pressure(const AcousticCoEvolverState* st, int n_rec,
                                     const float* rec_xyz, float* out_p);
SK_API int coevolver_get_string_velocity(const AcousticCoEvolverState* st, int string_idx,
                                         float* out_v, int out_len);
SK_API int coevolver_get_string_displacement(const AcousticCoEvolverState* st, int string_idx,
                                             float* out_d, int out_len);
SK_API int coevolver_get_string_position(const AcousticCoEvolverState* st, int string_idx,
                                         float* out_xyz, int out_len);
SK_API int coevolver_get_n_strings(const AcousticCoEvolverState* st);
SK_API int coevolver_get_n_pickups(const AcousticCoEvolverState* st);
SK_API int coevolver_get_n_mics(const AcousticCoEvolverState* st);
SK_API float coevolver_get_dt_audio(const AcousticCoEvolverState* st);
SK_API float coevolver_get_dt_fdtd(const AcousticCoEvolverState* st);

/* reference acoustic_coevolver.h:830-847, acoustic_coevolver.cpp:2698-2758.
 * Reset clears strings, neck, modal state, rings and counters and zeroes the
 * uniform body's fields; an AMR body is left as it is (the reference resets
 * only a uniform body, :2702). */
SK_API void coevolver_set_damping_scale(AcousticCoEvolverState* st, float scale);
SK_API int coevolver_reset(AcousticCoEvolverState* st);
SK_API int coevolver_set_face_fractions(AcousticCoEvolverState* st, const float* vx_frac,
                                        const float* vy_frac, const float* vz_frac);

/* reference acoustic_coevolver.h:872-897, acoustic_coevolver.cpp:2760-2797 */
SK_API int coevolver_get_surface_emission(const AcousticCoEvolverState* st, int n_surf,
                                          const float* surf_xyz, const float* surf_normals,
                                          float* out_P, float* out_vn);
SK_API int coevolver_sample_velocity(const AcousticCoEvolverState* st, int n_rec,
                                     const float* rec_xyz, float* out_vx, float* out_vy,
                                     float* out_vz);

/* ── Extensions (no reference counterpart) ─────────────────────────────── */

/* The body FDTD handle the co-evolver owns (borrowed; do not destroy).  Lets
 * a caller read any field through the fdtd_* ABI. */
SK_API AcousticFDTDState* coevolver_get_fdtd(AcousticCoEvolverState* st);

/* Body sub-steps per audio sample and string sub-steps per body step of one
 * string (the latter -1 for a bad index). */
SK_API int coevolver_get_substeps(const AcousticCoEvolverState* st);
SK_API int coevolver_get_string_substeps(const AcousticCoEvolverState* st, int string_idx);

/* Total samples produced since create / reset. */
SK_API int coevolver_get_samples_written(const AcousticCoEvolverState* st);

/* coevolver_step with the device time of the enqueued loop (ms, HIP events
 * around the whole block). */
SK_API int coevolver_step_timed(AcousticCoEvolverState* st, int n_samples, float* ms_total);

/* Raw state of one string: u[0], u[1], u_prev[0], u_prev[1] (n_nodes each)
 * in the string's local frame, then T_eff, saddle_w, saddle_w_prev,
 * bridge_w, bridge_w_prev and the string sub-step count; out_len must be
 * 4 * n_nodes + 6. */
SK_API int coevolver_get_string_state(const AcousticCoEvolverState* st, int string_idx,
                                      float* out, int out_len);

/* Shared neck mode: neck_w, neck_w_prev, the plate displacement it saw last. */
SK_API int coevolver_get_neck_state(const AcousticCoEvolverState* st, float* out3);

#ifdef __cplusplus
}
#endif

#endif /* AFDTD_COEVOLVER_H */
