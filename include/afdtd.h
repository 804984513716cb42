/*
 * afdtd.h — C ABI of the MI355X acoustic FDTD engine (libafdtd.so).
 *
 * Part 1 declares the 26 functions of the reference's uniform-grid FDTD ABI
 * with the same names, signatures, return codes and ownership rules, so a
 * caller that binds csrc/include/acoustic_fdtd.h (reference) can load this
 * library instead.  Each declaration cites the reference declaration it
 * replaces (csrc/include/acoustic_fdtd.h:line) and the definition whose
 * behaviour it reproduces (csrc/kernels/acoustic_fdtd.cpp:line).
 *
 * Part 2 declares extensions under new names (fdtd_create_ex, point sources,
 * receiver recording, energy audit, batched rooms, adjoint).  The reference
 * has no counterpart for them; the originals keep their exact semantics.
 *
 * Grid conventions (reference csrc/include/acoustic_fdtd.h:65-68):
 *   flat cell index = k + Nz*(j + Ny*i), z fastest.
 *   Vx packed (Nx-1)*Ny*Nz, index k + Nz*(j + Ny*i)
 *   Vy packed Nx*(Ny-1)*Nz, index k + Nz*(j + (Ny-1)*i)
 *   Vz packed Nx*Ny*(Nz-1), index k + (Nz-1)*(j + Ny*i)
 * (reference csrc/include/acoustic_fdtd.h:418-437).
 *
 * Device execution: every handle owns one HIP stream.  fdtd_step enqueues its
 * kernels and synchronises once before returning (it must report
 * FDTD_ERR_UNSTABLE).  Every get_* / sample_* call is stream-ordered and
 * synchronous on return.  Handles are not thread-safe (as in the reference).
 */
#ifndef AFDTD_H
#define AFDTD_H

#include <stddef.h>
#include <stdint.h>

#if defined(_WIN32)
#  define SK_API
#else
#  define SK_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Cell types — reference csrc/include/acoustic_fdtd.h:92-98 */
#define FDTD_AIR    ((uint8_t)0)
#define FDTD_WALL   ((uint8_t)1)
#define FDTD_PLATE  ((uint8_t)2)
#define FDTD_PML    ((uint8_t)3)

/* Error codes — reference csrc/include/acoustic_fdtd.h:102-105 */
#define FDTD_OK            0
#define FDTD_ERR_NULL     (-1)
#define FDTD_ERR_DIM      (-2)
#define FDTD_ERR_UNSTABLE (-3)
/* Extension error codes (never returned by the 26 reference functions). */
#define FDTD_ERR_DEVICE   (-4)   /* HIP runtime error                        */
#define FDTD_ERR_ARG      (-5)   /* unsupported option combination           */

typedef struct AcousticFDTDState AcousticFDTDState;

/* ═════════════════════ Part 1: the reference ABI ═══════════════════════ */

/* acoustic_fdtd.h:137 / acoustic_fdtd.cpp:327-768 */
SK_API AcousticFDTDState* fdtd_create(
    int Nx, int Ny, int Nz, float dx, float c, float rho_air,
    const uint8_t* cell_type, int plate_iz, const uint8_t* plate_active,
    float plate_mass_density, float plate_stiffness_D,
    float plate_alpha_M, float plate_beta_K, int n_pml);

/* acoustic_fdtd.h:153 / acoustic_fdtd.cpp:770-794 */
SK_API void fdtd_destroy(AcousticFDTDState* st);

/* acoustic_fdtd.h:170 / acoustic_fdtd.cpp:798-830 */
SK_API int fdtd_set_bridge_sources(AcousticFDTDState* st, int n_cells,
                                   const int* cell_indices, const float* weights);

/* acoustic_fdtd.h:184 / acoustic_fdtd.cpp:874-885 (disabled: returns -3) */
SK_API int fdtd_inject_bridge(AcousticFDTDState* st, float signal_val,
                              float signal_ddt, float force_scale);

/* acoustic_fdtd.h:207 / acoustic_fdtd.cpp:1699-1721 */
SK_API int fdtd_inject_bridge_drive(AcousticFDTDState* st,
                                    const float* cell_drives, float force_scale);

/* acoustic_fdtd.h:220 / acoustic_fdtd.cpp:1723-1742 */
SK_API int fdtd_inject_plate_force(AcousticFDTDState* st, int n_plate,
                                   const int* plate_indices, const float* weights,
                                   float total_force_N);

/* acoustic_fdtd.h:248 / acoustic_fdtd.cpp:1064-1077 */
SK_API int fdtd_set_face_fractions(AcousticFDTDState* st, const float* vx_frac,
                                   const float* vy_frac, const float* vz_frac);

/* acoustic_fdtd.h:257 / acoustic_fdtd.cpp:1744-1760 */
SK_API int fdtd_sample_plate_weighted(const AcousticFDTDState* st, int n_plate,
                                      const int* plate_indices, const float* weights,
                                      float* out_w);

/* acoustic_fdtd.h:281 / acoustic_fdtd.cpp:1567-1596 */
SK_API int fdtd_precompute_pressure_samplers(const AcousticFDTDState* st, int n_rec,
                                             const float* rec_xyz,
                                             int* out_idx8, float* out_wgt8);

/* acoustic_fdtd.h:296 / acoustic_fdtd.cpp:1647-1661 */
SK_API int fdtd_precompute_velocity_samplers(const AcousticFDTDState* st, int n_rec,
                                             const float* rec_xyz,
                                             int* out_vx_idx8, float* out_vx_wgt8,
                                             int* out_vy_idx8, float* out_vy_wgt8,
                                             int* out_vz_idx8, float* out_vz_wgt8);

/* acoustic_fdtd.h:308 / acoustic_fdtd.cpp:1663-1676 */
SK_API int fdtd_sample_pressure_precomputed(const AcousticFDTDState* st, int n_rec,
                                            const int* idx8, const float* wgt8,
                                            float* out_p);

/* No counterpart in the reference.  fdtd_sample_pressure_precomputed without
 * any device allocation, for callers that cannot allocate (tight serving
 * loops, harnesses that forbid hipMalloc): the stencils and outputs are
 * staged in a file-scope device array of the library (one per GPU, held
 * until the process ends) instead of buffers the handle allocates.  At most
 * FDTD_NOALLOC_MAX_RECEIVERS receivers per call (else FDTD_ERR_DIM); indices
 * outside [0, Nx*Ny*Nz) give FDTD_ERR_DIM.  Same result, bit for bit, as
 * fdtd_sample_pressure_precomputed.  All handles on one GPU share the array:
 * calls that overlap on one GPU (from several threads) race. */
#define FDTD_NOALLOC_MAX_RECEIVERS 65536
SK_API int fdtd_sample_pressure_precomputed_noalloc(const AcousticFDTDState* st, int n_rec,
                                                    const int* idx8, const float* wgt8,
                                                    float* out_p);

/* acoustic_fdtd.h:319 / acoustic_fdtd.cpp:1678-1697 */
SK_API int fdtd_sample_velocity_precomputed(const AcousticFDTDState* st, int n_rec,
                                            const int* vx_idx8, const float* vx_wgt8,
                                            const int* vy_idx8, const float* vy_wgt8,
                                            const int* vz_idx8, const float* vz_wgt8,
                                            float* out_vx, float* out_vy, float* out_vz);

/* acoustic_fdtd.h:339 / acoustic_fdtd.cpp:1272-1290 */
SK_API int fdtd_step(AcousticFDTDState* st, int n_steps);

/* acoustic_fdtd.h:342 / acoustic_fdtd.cpp:1294 */
SK_API int fdtd_get_step_count(const AcousticFDTDState* st);

/* acoustic_fdtd.h:345 / acoustic_fdtd.cpp:1295 */
SK_API float fdtd_get_dt(const AcousticFDTDState* st);

/* acoustic_fdtd.h:348 / acoustic_fdtd.cpp:1297-1310 */
SK_API int fdtd_reset(AcousticFDTDState* st);

/* acoustic_fdtd.h:359 / acoustic_fdtd.cpp:1314-1321 */
SK_API int fdtd_get_pressure_field(const AcousticFDTDState* st, float* out, int out_len);

/* acoustic_fdtd.h:372 / acoustic_fdtd.cpp:1323-1330 */
SK_API int fdtd_get_plate_displacement(const AcousticFDTDState* st, float* out,
                                       int out_len);

/* acoustic_fdtd.h:394 / acoustic_fdtd.cpp:1332-1353 */
SK_API int fdtd_sample_plate_displacement_batch(const AcousticFDTDState* st,
                                                int n_points, const int* saddle_idx4,
                                                const float* saddle_wgt4, float* out_w);

/* acoustic_fdtd.h:411 / acoustic_fdtd.cpp:1355-1395 */
SK_API int fdtd_sample_pressure(const AcousticFDTDState* st, int n_rec,
                                const float* rec_xyz, float* out_p);

/* acoustic_fdtd.h:440 / acoustic_fdtd.cpp:1399-1406 */
SK_API void fdtd_get_velocity_dims(const AcousticFDTDState* st,
                                   int* Nvx, int* Nvy, int* Nvz);

/* acoustic_fdtd.h:448-456 / acoustic_fdtd.cpp:1408-1430 */
SK_API int fdtd_get_velocity_x(const AcousticFDTDState* st, float* out, int out_len);
SK_API int fdtd_get_velocity_y(const AcousticFDTDState* st, float* out, int out_len);
SK_API int fdtd_get_velocity_z(const AcousticFDTDState* st, float* out, int out_len);

/* acoustic_fdtd.h:475 / acoustic_fdtd.cpp:1523-1538 */
SK_API int fdtd_sample_velocity(const AcousticFDTDState* st, int n_rec,
                                const float* rec_xyz,
                                float* out_vx, float* out_vy, float* out_vz);

/* acoustic_fdtd.h:499 / acoustic_fdtd.cpp:1540-1563 */
SK_API int fdtd_get_surface_emission(const AcousticFDTDState* st, int n_surf,
                                     const float* surf_xyz, const float* surf_normals,
                                     float* out_P, float* out_vn);

/* ═════════════════════ Part 2: extensions (no reference counterpart) ════ */

/* Stencil order.  4 = the reference's Fornberg 9/8, -1/24 staggered stencil
 * with 2nd-order edge faces (acoustic_fdtd.cpp:889-890, :1107-1112).
 * 2 = the Yee stencil of the reference's AMR order-2 path
 * (acoustic_amr.cpp:1146-1154) on the uniform grid. */
#define FDTD_ORDER_4 4
#define FDTD_ORDER_2 2

/* Domain-boundary model. */
#define FDTD_BOUNDARY_RIGID_PML  0  /* reference: rigid domain edges + n_pml PML   */
#define FDTD_BOUNDARY_IMPEDANCE  1  /* locally reacting walls, reflection refl[6]  */

/* Field storage type (arithmetic is always FP32). */
#define FDTD_STORE_F32  0
#define FDTD_STORE_BF16 1
#define FDTD_STORE_F16  2
/* Block floating point: int16 mantissas plus one power-of-two exponent per
 * 64-cell z-row segment (absolute error <= 2^-16 of the segment's largest
 * magnitude).  Fused sweep only: no walls, plate cells, face fractions or
 * active plate. */
#define FDTD_STORE_BFP16 3
/* FP16 with a per-row power-of-two scale: FP16 values plus one exponent per
 * 64-cell z-row segment (the layout and exponent rule of BFP16; relative
 * error <= 2^-11 per value down to 2^-28 of the segment's largest
 * magnitude).  Fused sweep only, as BFP16. */
#define FDTD_STORE_F16S 4

/* Kernel family selection. */
#define FDTD_KERNEL_AUTO    0
#define FDTD_KERNEL_GENERIC 1  /* separate velocity / pressure sweeps, all features */
#define FDTD_KERNEL_FUSED   2  /* fused 2.5-D plane-marching sweep                   */
#define FDTD_KERNEL_FUSED2  3  /* two leapfrog steps per sweep (order 2)            */

typedef struct FdtdOptions {
    int          struct_size;    /* sizeof(FdtdOptions)                          */
    int          stencil_order;  /* FDTD_ORDER_4 (default) or FDTD_ORDER_2        */
    float        courant;        /* <= 0: reference dt = 0.77*dx/(c*sqrt3);
                                    > 0: dt = courant*dx/c_max                   */
    int          boundary;       /* FDTD_BOUNDARY_*                               */
    float        refl[6];        /* reflection coefficient per domain side
                                    (x-, x+, y-, y+, z-, z+), impedance only       */
    const float* rho;            /* per-cell density (N) or NULL = rho_air        */
    const float* bulk;           /* per-cell bulk modulus (N) or NULL = rho*c^2   */
    int          storage;        /* FDTD_STORE_*                                  */
    int          kernel;         /* FDTD_KERNEL_*                                 */
    int          tile_y;         /* fused-kernel tile rows (0 = autotuned)        */
    int          seg_x;          /* fused-kernel x-segment length (0 = autotuned) */
} FdtdOptions;

SK_API void fdtd_default_options(FdtdOptions* opt);

/* fdtd_create with the options above.  opt == NULL gives fdtd_create. */
SK_API AcousticFDTDState* fdtd_create_ex(
    int Nx, int Ny, int Nz, float dx, float c, float rho_air,
    const uint8_t* cell_type, int plate_iz, const uint8_t* plate_active,
    float plate_mass_density, float plate_stiffness_D,
    float plate_alpha_M, float plate_beta_K, int n_pml,
    const FdtdOptions* opt);

/* Point pressure sources: after the pressure update of the step with global
 * index n (0-based step_count before the step), P[cell_idx[s]] +=
 * signals[s*n_samples + n] for n < n_samples (additive, as
 * acoustic_amr.cpp:1227-1244).  Sources on solid cells are rejected. */
SK_API int fdtd_set_point_sources(AcousticFDTDState* st, int n_src,
                                  const int* cell_idx, const float* signals,
                                  int n_samples);

/* Receivers at grid coordinates, gathered on the device after every step with
 * the 8-tap trilinear stencil of fdtd_precompute_pressure_samplers. */
SK_API int fdtd_set_receivers(AcousticFDTDState* st, int n_rec, const float* rec_xyz);
/* Start (or restart) recording up to max_steps steps of receiver traces. */
SK_API int fdtd_record_receivers(AcousticFDTDState* st, int max_steps);
/* Copy recorded traces (n_recorded x n_rec, row = step). */
SK_API int fdtd_get_receiver_traces(const AcousticFDTDState* st, float* out,
                                    int out_len, int* n_recorded);

/* FP64 energy audit: out[0] = sum p^2/(2K) dV, out[1] = sum rho_f v^2/2 dV
 * (velocities at the current half step). */
SK_API int fdtd_energy(const AcousticFDTDState* st, double* out2);

/* Upload a full state (P then packed Vx, Vy, Vz) — used by tests and the
 * checkpoint/recompute adjoint. */
SK_API int fdtd_set_state(AcousticFDTDState* st, const float* P, const float* vx,
                          const float* vy, const float* vz);

/* Introspection: which kernel family runs, dt, stencil order. */
SK_API int fdtd_get_kernel_family(const AcousticFDTDState* st);
/* Extension (no counterpart in the reference): steps one sweep advances, 2
 * for the two-step sweeps (rooms, and order-4 FP32 grids with walls, face
 * fractions or an active plate), else 1. */
SK_API int fdtd_get_sweep_steps(const AcousticFDTDState* st);
SK_API const char* fdtd_last_error(void);

/* Timing helper: run n_steps with HIP events around the step kernels and
 * return the average duration of one launch of the dominant (stencil) kernel
 * in *ms_per_launch and the wall time of the whole call in *ms_total. */
SK_API int fdtd_step_timed(AcousticFDTDState* st, int n_steps,
                           float* ms_total, float* ms_stencil_avg);

/* Device-bandwidth probe (no reference counterpart): time a 16-byte-per-lane
 * device-to-device copy of nbytes over `reps` launches; *ms_per_copy is the
 * average launch time.  Reported as the measured copy ceiling. */
SK_API int fdtd_copy_calibration(size_t nbytes, int reps, float* ms_per_copy);

/* Streaming ceiling of a stencil's byte mix (no reference counterpart): a
 * kernel that reads n_read (1..8) and writes n_write (1..8) FP32 arrays of
 * n_cells elements once per pass, with no stencil.  *ms_per_pass is the
 * HIP-event time per pass over `reps` passes.  Device buffers are allocated
 * and freed inside the call. */
SK_API int fdtd_stream_calibration(size_t n_cells, int n_read, int n_write, int reps,
                                   float* ms_per_pass);

/* ── Batched rooms (no reference counterpart) ─────────────────────────────
 * n_rooms independent rooms of uneven size advanced by ONE fused launch per
 * step (plus one source/receiver launch), the whole n-step sequence captured
 * in a hipGraph.  dims: n_rooms x (Nx, Ny, Nz); courant: per-room Courant
 * number (dt = courant*dx/c); refl: n_rooms x 6 reflection coefficients
 * (impedance walls, order 2) or NULL for rigid walls. */
typedef struct FdtdBatch FdtdBatch;
SK_API FdtdBatch* fdtd_batch_create(int n_rooms, const int* dims, float dx, float c,
                                    float rho_air, const float* courant, const float* refl,
                                    int stencil_order, int storage, int tile_y, int seg_x);
SK_API void fdtd_batch_destroy(FdtdBatch* b);
SK_API int fdtd_batch_cells(const FdtdBatch* b, long long* total, int* n_tiles);
/* Autotuner (before the first step): times every sweep (single-step; for
 * order 2 also the two-step sweep), tile height (8, 16, 32 rows; 8, 16 for
 * the two-step sweep), x-segment (16..128 planes) and dispatch order (room
 * order / largest rooms first) with trial_steps graph-captured steps of this
 * batch, keeps the fastest, and leaves the batch at step 0 with zero fields.
 * Results do not depend on the choice (same arithmetic per cell). */
SK_API int fdtd_batch_autotune(FdtdBatch* b, int trial_steps, int* best_tile_y, int* best_seg,
                               int* best_schedule, float* best_ms_per_step);
/* Sweep of the batch: FDTD_KERNEL_AUTO (default: two steps per launch with the
 * two-step sweep where it applies, i.e. order 2 and tile height 8 or 16),
 * FDTD_KERNEL_FUSED (single steps) or FDTD_KERNEL_FUSED2 (two-step sweep
 * required).  No reference counterpart (extension). */
SK_API int fdtd_batch_set_kernel(FdtdBatch* b, int family);
/* The sweep the next steps use: FDTD_KERNEL_FUSED2 or FDTD_KERNEL_FUSED. */
SK_API int fdtd_batch_get_kernel_family(FdtdBatch* b);
SK_API int fdtd_batch_set_sources(FdtdBatch* b, const int* cell_idx, const float* signals,
                                  int n_samples);
SK_API int fdtd_batch_set_receivers(FdtdBatch* b, int per_room, const float* xyz);
SK_API int fdtd_batch_record(FdtdBatch* b, int max_steps);
SK_API int fdtd_batch_step(FdtdBatch* b, int n_steps, int use_graph, float* ms_total);
SK_API int fdtd_batch_get_traces(const FdtdBatch* b, float* out, int len, int* n_recorded);
SK_API int fdtd_batch_get_pressure(const FdtdBatch* b, int room, float* out, int len);
SK_API int fdtd_batch_get_step_count(const FdtdBatch* b);

/* ── Adjoint (no reference counterpart) ──────────────────────────────────
 * Gradients of receiver traces of the order-2 heterogeneous step with
 * impedance walls with respect to per-cell m = log rho, l = log K and the
 * per-boundary-cell admittance sum eta.  precision_bytes 4 or 8 selects the
 * REAL type of every array; parameters, traces and gradients are DEVICE
 * arrays (e.g. torch tensors), sources/receivers are host arrays.  The
 * forward stores (P, V) checkpoints every checkpoint_every steps; the
 * backward recomputes each segment and runs the time-reversed adjoint. */
typedef struct FdtdAdjoint FdtdAdjoint;
SK_API FdtdAdjoint* fdtd_adjoint_create(int Nx, int Ny, int Nz, double dtdx,
                                        int precision_bytes, int checkpoint_every);
SK_API void fdtd_adjoint_destroy(FdtdAdjoint* a);
SK_API int fdtd_adjoint_set_sources(FdtdAdjoint* a, int n_src, const int* idx,
                                    const void* signals, int n_samples);
SK_API int fdtd_adjoint_set_receivers(FdtdAdjoint* a, int n_rec, const int* idx8,
                                      const void* wgt8);
SK_API int fdtd_adjoint_forward(FdtdAdjoint* a, const void* m, const void* l,
                                const void* eta, int n_steps, void* traces);
SK_API int fdtd_adjoint_backward(FdtdAdjoint* a, const void* grad_traces, void* grad_m,
                                 void* grad_l, void* grad_eta);

/* ── Geometry input formats (extension) ───────────────────────────────────
 * The body voxeliser and the cut-cell face fractions that produce the
 * cell_type / plate_active arrays of fdtd_create and the face fractions of
 * fdtd_set_face_fractions, computed on the GPU.  The reference computes them
 * in Python above its C ABI: voxelise_guitar_body
 * (acoustic_fdtd_bridge.py:255-359, non-AMR branch) and
 * compute_cut_cell_fractions (acoustic_fdtd_bridge.py:456-552); the results
 * are equal element for element.  outline_xy: n_pts (x, y) float pairs, the
 * closed outline polygon, 3 <= n_pts <= 4096 (else FDTD_ERR_DIM).  The grid (dimensions, offsets g*_min, spacing dx,
 * body extent z_min..z_max and the back / plate layers iz_back / iz_top) is
 * the one voxelise_guitar_body derives from the outline's bounding box.
 * Layouts: cell_type Nx*Ny*Nz (k fastest), plate_active Nx*Ny, fractions in
 * the packed staggered layouts of fdtd_get_velocity_x/y/z.  The plain forms
 * take host memory; the _device forms take device memory and run on `stream`
 * (a hipStream_t of this library's HIP runtime, NULL = its default stream),
 * and when kernel_ms is not NULL store the kernels' time in milliseconds
 * (HIP events on that stream).  Every form returns when its results are
 * written. */
SK_API int fdtd_voxelise_body(const float* outline_xy, int n_pts, int Nx, int Ny, int Nz,
                              double gx_min, double gy_min, double gz_min, double dx,
                              double z_min, double z_max, int iz_back, int iz_top,
                              unsigned char* cell_type, unsigned char* plate_active);
SK_API int fdtd_voxelise_body_device(const float* outline_xy, int n_pts, int Nx, int Ny, int Nz,
                                     double gx_min, double gy_min, double gz_min, double dx,
                                     double z_min, double z_max, int iz_back, int iz_top,
                                     unsigned char* cell_type, unsigned char* plate_active,
                                     void* stream, float* kernel_ms);
SK_API int fdtd_cut_cell_fractions(const unsigned char* cell_type,
                                   const unsigned char* plate_active, const float* outline_xy,
                                   int n_pts, int Nx, int Ny, int Nz, double gx_min,
                                   double gy_min, double gz_min, double dx, int plate_iz,
                                   int iz_back, float* vx_frac, float* vy_frac, float* vz_frac);
SK_API int fdtd_cut_cell_fractions_device(const unsigned char* cell_type,
                                          const unsigned char* plate_active,
                                          const float* outline_xy, int n_pts, int Nx, int Ny,
                                          int Nz, double gx_min, double gy_min, double gz_min,
                                          double dx, int plate_iz, int iz_back, float* vx_frac,
                                          float* vy_frac, float* vz_frac, void* stream,
                                          float* kernel_ms);

#ifdef __cplusplus
}
#endif
#endif /* AFDTD_H */
