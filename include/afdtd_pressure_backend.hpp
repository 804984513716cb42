/*
 * afdtd_pressure_backend.hpp — the reference's pressure-backend interface
 * (csrc/include/acoustic_pressure_backend.h:36-201) over libafdtd.so, so a
 * C++ caller above it (the co-evolver) can step either domain on the GPU:
 *
 *   UniformPressureBackend — wraps an AcousticFDTDState (include/afdtd.h),
 *     reference csrc/kernels/acoustic_pressure_backend.cpp:64-258;
 *   AMR backend — wraps an AcousticAMRState (include/afdtd_amr.h), built from
 *     an AMRCoevolverDescriptor by amr_pressure_backend_create, reference
 *     csrc/kernels/acoustic_pressure_backend.cpp:273-569.
 *
 * The interface, the descriptor layout (reference
 * csrc/include/acoustic_coevolver.h:331-397), the creation error codes and
 * messages are the reference's; the C++ symbols have C++ linkage, as there.
 */
#ifndef AFDTD_PRESSURE_BACKEND_HPP
#define AFDTD_PRESSURE_BACKEND_HPP

#include <cstdint>

#include "afdtd.h"
#include "afdtd_amr.h"
#include "afdtd_coevolver.h"

#define AFDTD_BACKEND_API __attribute__((visibility("default")))

/* AMRCoevolverDescriptor: include/afdtd_coevolver.h (reference
 * acoustic_coevolver.h:331-397) */

/* acoustic_pressure_backend.h:36-201 */
struct AFDTD_BACKEND_API IPressureBackend {
    virtual ~IPressureBackend() = default;
    virtual double get_dt() const = 0;
    virtual int step(int n_steps) = 0;
    virtual int reset() = 0;
    virtual int inject_bridge_drive(const float* cell_drives, int n_cells, float force_scale) = 0;
    virtual int inject_plate_force(int n_plate, const int* plate_indices, const float* weights,
                                   float total_force_N) = 0;
    virtual int sample_plate_displacement_batch(int n_points, const int* idx4, const float* wgt4,
                                                float* out_w) = 0;
    virtual int sample_plate_weighted(int n_plate, const int* plate_indices, const float* weights,
                                      float* out_w) = 0;
    virtual int setup_mic_samplers(int n_mics, const float* pos_xyz, const float* axis_xyz,
                                   int need_veloc) = 0;
    virtual int sample_mics(int n_mics, float* out_p, float* out_vx, float* out_vy,
                            float* out_vz) = 0;
    virtual int get_pressure_field_size() const = 0;
    virtual int get_pressure_field_flat(float* out, int out_len) = 0;
    virtual int get_cell_centers(float* out_xyz, int n_cells_3) {
        (void)out_xyz;
        (void)n_cells_3;
        return SK_ERR_DIM_MISMATCH;
    }
    virtual int get_pressure_field_uniform(int Nx, int Ny, int Nz, const double bmin[3],
                                           const double bmax[3], float* out, int out_len) = 0;
    virtual int get_plate_dims(int* out_Nx, int* out_Ny, int* out_count) const = 0;
    virtual int get_plate_displacement_size() const = 0;
    virtual int get_plate_displacement(float* out, int out_len) = 0;
};

/* acoustic_pressure_backend.h:216-289: owns the handle (destroyed with the
 * backend); mic samplers are the 8-tap pressure / staggered velocity
 * stencils of the uniform engine. */
class AFDTD_BACKEND_API UniformPressureBackend : public IPressureBackend {
public:
    UniformPressureBackend(AcousticFDTDState* fdtd, int Nx, int Ny, int Nz, float dx,
                           const float origin[3], float rho_air, float c_sound);
    ~UniformPressureBackend() override;
    double get_dt() const override;
    int step(int n_steps) override;
    int reset() override;
    int inject_bridge_drive(const float* cell_drives, int n_cells, float force_scale) override;
    int inject_plate_force(int n_plate, const int* plate_indices, const float* weights,
                           float total_force_N) override;
    int sample_plate_displacement_batch(int n_points, const int* idx4, const float* wgt4,
                                        float* out_w) override;
    int sample_plate_weighted(int n_plate, const int* plate_indices, const float* weights,
                              float* out_w) override;
    int setup_mic_samplers(int n_mics, const float* pos_xyz, const float* axis_xyz,
                           int need_veloc) override;
    int sample_mics(int n_mics, float* out_p, float* out_vx, float* out_vy,
                    float* out_vz) override;
    int get_pressure_field_size() const override;
    int get_pressure_field_flat(float* out, int out_len) override;
    int get_pressure_field_uniform(int Nx, int Ny, int Nz, const double bmin[3],
                                   const double bmax[3], float* out, int out_len) override;
    int get_plate_dims(int* out_Nx, int* out_Ny, int* out_count) const override;
    int get_plate_displacement_size() const override;
    int get_plate_displacement(float* out, int out_len) override;
    AcousticFDTDState* fdtd() const { return fdtd_; }

private:
    AcousticFDTDState* fdtd_ = nullptr;
    int Nx_ = 0, Ny_ = 0, Nz_ = 0;                /* dx, origin, rho, c: not needed */
    int n_mics_ = 0, need_v_ = 0;
    int* idx_[4] = {nullptr, nullptr, nullptr, nullptr};     /* P, Vx, Vy, Vz: 8 per mic */
    float* wgt_[4] = {nullptr, nullptr, nullptr, nullptr};
};

/* acoustic_pressure_backend.h:291-307 */
enum AmrPressureBackendCreateError {
    AMR_BACKEND_CREATE_OK = 0,
    AMR_BACKEND_DESC_NULL,
    AMR_BACKEND_DESC_INVALID_REQUIRED_FIELD,
    AMR_BACKEND_DESC_INVALID_PLATE,
    AMR_BACKEND_AMR_CREATE_FAILED,
    AMR_BACKEND_AMR_SETUP_PLATE_FAILED,
    AMR_BACKEND_BRIDGE_SOURCE_FAILED,
    AMR_BACKEND_NECK_SOURCE_FAILED,
    AMR_BACKEND_BORDER_SETUP_FAILED,
    AMR_BACKEND_NEW_FAILED,
};

AFDTD_BACKEND_API const char* amr_pressure_backend_last_error_code(void);
AFDTD_BACKEND_API const char* amr_pressure_backend_last_error_message(void);

/* Validates the descriptor as the reference does (required fields, plate,
 * bridge mapping, dt sanity), then amr_create, amr_setup_plate, bridge and
 * optional neck sources, and the border condition when n_pml > 0; NULL on
 * any failure with the reason in the two getters above.  step() reports a
 * divergence as FDTD_ERR_UNSTABLE, as the co-evolver expects. */
AFDTD_BACKEND_API IPressureBackend* amr_pressure_backend_create(const AMRCoevolverDescriptor* desc);

#endif /* AFDTD_PRESSURE_BACKEND_HPP */
