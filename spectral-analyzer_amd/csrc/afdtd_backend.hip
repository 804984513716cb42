// afdtd_backend.hip — host code only: the pressure-backend interface of
// include/afdtd_pressure_backend.hpp over the uniform (fdtd_*) and AMR
// (amr_*) entry points of this library.  Both backends are thin: every
// call forwards to the C ABI, whose kernels do the work on the handle's
// stream.  Return-code mapping, descriptor validation order and error
// messages follow csrc/kernels/acoustic_pressure_backend.cpp.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <cstddef>
#include <new>
#include <string>

#include "../../include/afdtd_pressure_backend.hpp"
#include "afdtd_internal.h"

// ── UniformPressureBackend (acoustic_pressure_backend.cpp:64-258) ────────

UniformPressureBackend::UniformPressureBackend(AcousticFDTDState* fdtd, int Nx, int Ny, int Nz,
                                               float dx, const float origin[3], float rho_air,
                                               float c_sound)
    : fdtd_(fdtd), Nx_(Nx), Ny_(Ny), Nz_(Nz) {
    (void)dx;                                        // kept in the reference, never read
    (void)origin;
    (void)rho_air;
    (void)c_sound;
}

UniformPressureBackend::~UniformPressureBackend() {
    fdtd_destroy(fdtd_);
    for (int k = 0; k < 4; ++k) {
        free(idx_[k]);
        free(wgt_[k]);
    }
}

double UniformPressureBackend::get_dt() const { return fdtd_ ? (double)fdtd_get_dt(fdtd_) : 0.0; }

int UniformPressureBackend::step(int n_steps) {
    if (!fdtd_) return SK_ERR_NULL_STATE;
    const int rc = fdtd_step(fdtd_, n_steps);
    if (rc == FDTD_ERR_UNSTABLE) return FDTD_ERR_UNSTABLE;
    return rc == FDTD_OK ? SK_OK : SK_ERR_DIM_MISMATCH;
}

int UniformPressureBackend::reset() {
    if (!fdtd_) return SK_ERR_NULL_STATE;
    fdtd_reset(fdtd_);
    return SK_OK;
}

int UniformPressureBackend::inject_bridge_drive(const float* drives, int n_cells, float scale) {
    (void)n_cells;                                   // the handle knows its bridge cells
    if (!fdtd_ || !drives) return SK_ERR_NULL_STATE;
    return fdtd_inject_bridge_drive(fdtd_, drives, scale) == FDTD_OK ? SK_OK : SK_ERR_DIM_MISMATCH;
}

int UniformPressureBackend::inject_plate_force(int n, const int* idx, const float* w, float total) {
    if (!fdtd_) return SK_ERR_NULL_STATE;
    if (n <= 0 || !idx || !w) return SK_ERR_DIM_MISMATCH;
    return fdtd_inject_plate_force(fdtd_, n, idx, w, total) == FDTD_OK ? SK_OK : SK_ERR_DIM_MISMATCH;
}

int UniformPressureBackend::sample_plate_displacement_batch(int n, const int* idx4,
                                                            const float* wgt4, float* out) {
    if (!fdtd_ || !idx4 || !wgt4 || !out) return SK_ERR_NULL_STATE;
    return fdtd_sample_plate_displacement_batch(fdtd_, n, idx4, wgt4, out) == FDTD_OK
               ? SK_OK
               : SK_ERR_DIM_MISMATCH;
}

int UniformPressureBackend::sample_plate_weighted(int n, const int* idx, const float* w,
                                                  float* out) {
    if (!fdtd_ || !idx || !w || !out) return SK_ERR_NULL_STATE;
    return fdtd_sample_plate_weighted(fdtd_, n, idx, w, out) == FDTD_OK ? SK_OK
                                                                        : SK_ERR_DIM_MISMATCH;
}

int UniformPressureBackend::setup_mic_samplers(int n_mics, const float* pos, const float* axis,
                                               int need_v) {
    if (!fdtd_ || !pos || !axis) return SK_ERR_NULL_STATE;
    if (n_mics <= 0) return SK_ERR_DIM_MISMATCH;
    for (int k = 0; k < 4; ++k) {
        free(idx_[k]);
        free(wgt_[k]);
        idx_[k] = nullptr;
        wgt_[k] = nullptr;
    }
    n_mics_ = n_mics;
    need_v_ = need_v;
    for (int k = 0; k < (need_v ? 4 : 1); ++k) {
        idx_[k] = (int*)calloc((size_t)n_mics * 8, sizeof(int));
        wgt_[k] = (float*)calloc((size_t)n_mics * 8, sizeof(float));
        if (!idx_[k] || !wgt_[k]) return SK_ERR_NULL_STATE;
    }
    if (fdtd_precompute_pressure_samplers(fdtd_, n_mics, pos, idx_[0], wgt_[0]) != FDTD_OK)
        return SK_ERR_DIM_MISMATCH;
    if (need_v && fdtd_precompute_velocity_samplers(fdtd_, n_mics, pos, idx_[1], wgt_[1], idx_[2],
                                                    wgt_[2], idx_[3], wgt_[3]) != FDTD_OK)
        return SK_ERR_DIM_MISMATCH;
    return SK_OK;
}

int UniformPressureBackend::sample_mics(int n_mics, float* op, float* ovx, float* ovy,
                                        float* ovz) {
    if (!fdtd_ || !op) return SK_ERR_NULL_STATE;
    if (n_mics != n_mics_) return SK_ERR_DIM_MISMATCH;
    if (fdtd_sample_pressure_precomputed(fdtd_, n_mics, idx_[0], wgt_[0], op) != FDTD_OK)
        return SK_ERR_DIM_MISMATCH;
    if (need_v_ && ovx && ovy && ovz &&
        fdtd_sample_velocity_precomputed(fdtd_, n_mics, idx_[1], wgt_[1], idx_[2], wgt_[2],
                                         idx_[3], wgt_[3], ovx, ovy, ovz) != FDTD_OK)
        return SK_ERR_DIM_MISMATCH;
    return SK_OK;
}

int UniformPressureBackend::get_pressure_field_size() const { return Nx_ * Ny_ * Nz_; }

int UniformPressureBackend::get_pressure_field_flat(float* out, int out_len) {
    if (!fdtd_ || !out) return SK_ERR_NULL_STATE;
    if (out_len != Nx_ * Ny_ * Nz_) return SK_ERR_DIM_MISMATCH;
    return fdtd_get_pressure_field(fdtd_, out, out_len) == FDTD_OK ? SK_OK : SK_ERR_DIM_MISMATCH;
}

int UniformPressureBackend::get_pressure_field_uniform(int Nx, int Ny, int Nz, const double*,
                                                       const double*, float* out, int out_len) {
    if (Nx != Nx_ || Ny != Ny_ || Nz != Nz_) return SK_ERR_DIM_MISMATCH;  // already Cartesian
    return get_pressure_field_flat(out, out_len);
}

int UniformPressureBackend::get_plate_displacement_size() const { return Nx_ * Ny_; }

int UniformPressureBackend::get_plate_dims(int* nx, int* ny, int* count) const {
    if (!nx || !ny || !count) return SK_ERR_NULL_STATE;
    *nx = Nx_;
    *ny = Ny_;
    *count = Nx_ * Ny_;
    return SK_OK;
}

int UniformPressureBackend::get_plate_displacement(float* out, int out_len) {
    if (!fdtd_ || !out) return SK_ERR_NULL_STATE;
    if (out_len != Nx_ * Ny_) return SK_ERR_DIM_MISMATCH;
    return fdtd_get_plate_displacement(fdtd_, out, out_len) == FDTD_OK ? SK_OK
                                                                       : SK_ERR_DIM_MISMATCH;
}

// ── AMR backend (acoustic_pressure_backend.cpp:273-569) ──────────────────

namespace {

thread_local AmrPressureBackendCreateError g_code = AMR_BACKEND_CREATE_OK;
thread_local std::string g_msg;

const char* code_name(AmrPressureBackendCreateError c) {
    switch (c) {
        case AMR_BACKEND_CREATE_OK: return "OK";
        case AMR_BACKEND_DESC_NULL: return "DESC_NULL";
        case AMR_BACKEND_DESC_INVALID_REQUIRED_FIELD: return "DESC_INVALID_REQUIRED_FIELD";
        case AMR_BACKEND_DESC_INVALID_PLATE: return "DESC_INVALID_PLATE";
        case AMR_BACKEND_AMR_CREATE_FAILED: return "AMR_CREATE_FAILED";
        case AMR_BACKEND_AMR_SETUP_PLATE_FAILED: return "AMR_SETUP_PLATE_FAILED";
        case AMR_BACKEND_BRIDGE_SOURCE_FAILED: return "BRIDGE_SOURCE_FAILED";
        case AMR_BACKEND_NECK_SOURCE_FAILED: return "NECK_SOURCE_FAILED";
        case AMR_BACKEND_BORDER_SETUP_FAILED: return "BORDER_SETUP_FAILED";
        case AMR_BACKEND_NEW_FAILED: return "BACKEND_NEW_FAILED";
    }
    return "UNKNOWN";
}

std::nullptr_t fail(AmrPressureBackendCreateError c, const std::string& m,
                    AcousticAMRState* amr = nullptr) {
    g_code = c;
    g_msg = m;
    amr_destroy(amr);
    return nullptr;
}

class AMRPressureBackend final : public IPressureBackend {
public:
    AMRPressureBackend(AcousticAMRState* amr, int n_cells, int pNx, int pNy)
        : amr_(amr), n_cells_(n_cells), pNx_(pNx), pNy_(pNy) {}
    ~AMRPressureBackend() override { amr_destroy(amr_); }
    double get_dt() const override { return amr_get_dt(amr_); }
    int step(int n) override {
        const int rc = amr_step(amr_, n);
        return rc == SK_ERR_DIVERGED ? FDTD_ERR_UNSTABLE : rc;
    }
    int reset() override { return amr_reset(amr_); }
    int inject_bridge_drive(const float* d, int n, float s) override {
        return amr_inject_bridge_drive(amr_, d, n, s);
    }
    int inject_plate_force(int n, const int* i, const float* w, float t) override {
        return amr_inject_plate_force(amr_, n, i, w, t);
    }
    int sample_plate_displacement_batch(int n, const int* i4, const float* w4, float* o) override {
        return amr_sample_plate_displacement_batch(amr_, n, i4, w4, o);
    }
    int sample_plate_weighted(int n, const int* i, const float* w, float* o) override {
        return amr_sample_plate_weighted(amr_, n, i, w, o);
    }
    int setup_mic_samplers(int n, const float* p, const float* a, int v) override {
        return amr_setup_mic_samplers(amr_, n, p, a, v);
    }
    int sample_mics(int n, float* p, float* vx, float* vy, float* vz) override {
        return amr_sample_mics(amr_, n, p, vx, vy, vz);
    }
    int get_pressure_field_size() const override { return n_cells_; }
    int get_pressure_field_flat(float* out, int len) override {
        return amr_get_pressure(amr_, out, len);
    }
    int get_cell_centers(float* out, int n3) override { return amr_get_cell_centers(amr_, out, n3); }
    int get_pressure_field_uniform(int Nx, int Ny, int Nz, const double bmin[3],
                                   const double bmax[3], float* out, int len) override {
        if (len != Nx * Ny * Nz) return SK_ERR_DIM_MISMATCH;
        return amr_scatter_pressure_uniform(amr_, Nx, Ny, Nz, bmin, bmax, out);
    }
    int get_plate_displacement_size() const override { return pNx_ * pNy_; }
    int get_plate_dims(int* nx, int* ny, int* count) const override {
        if (!nx || !ny || !count) return SK_ERR_NULL_STATE;
        *nx = pNx_;
        *ny = pNy_;
        *count = pNx_ * pNy_;
        return SK_OK;
    }
    int get_plate_displacement(float* out, int len) override {
        return amr_get_plate_displacement(amr_, out, len);
    }

private:
    AcousticAMRState* amr_;
    int n_cells_, pNx_, pNy_;
};

}  // namespace

const char* amr_pressure_backend_last_error_code(void) { return code_name(g_code); }
const char* amr_pressure_backend_last_error_message(void) { return g_msg.c_str(); }

AcousticAMRState* amr_state_from_descriptor(const AMRCoevolverDescriptor* d) {
    g_code = AMR_BACKEND_CREATE_OK;
    g_msg.clear();
    if (!d) return fail(AMR_BACKEND_DESC_NULL, "AMRCoevolverDescriptor is null");
    const auto REQ = AMR_BACKEND_DESC_INVALID_REQUIRED_FIELD;
    const auto PLT = AMR_BACKEND_DESC_INVALID_PLATE;
    if (d->n_cells <= 0 || d->n_faces <= 0) return fail(REQ, "n_cells and n_faces must be > 0");
    if (!d->cell_centers || !d->cell_volumes || !d->open_volume_frac || !d->cell_types ||
        !d->cell_levels)
        return fail(REQ, "missing one or more required cell arrays");
    if (!d->face_cell_neg || !d->face_cell_pos || !d->face_area || !d->face_open_frac ||
        !d->face_distance)
        return fail(REQ, "missing one or more required face arrays");
    if (d->c <= 0.0 || d->rho_air <= 0.0 || d->min_dx <= 0.0)
        return fail(REQ, "c, rho_air, and min_dx must be > 0");
    if (d->n_bridge_plate <= 0 || !d->bridge_plate_idx || !d->bridge_plate_wgt)
        return fail(REQ, "bridge plate mapping is missing or empty");
    if (d->plate_Nx <= 0 || d->plate_Ny <= 0 || d->plate_dx <= 0.0f)
        return fail(PLT, "plate_Nx/plate_Ny/plate_dx must be valid");
    if (!d->plate_active || !d->plate_active_flat_idx)
        return fail(PLT, "plate_active or plate_active_flat_idx is missing");
    if (!d->plate_face_above_starts || !d->plate_face_below_starts)
        return fail(PLT, "plate face start arrays are missing");
    if (!d->plate_cell_above || !d->plate_cell_below)
        return fail(PLT, "plate cell mapping arrays are missing");
    if (d->plate_mass_density <= 0.0f || d->plate_stiffness_D <= 0.0f)
        return fail(PLT, "plate mass density and stiffness must be > 0");
    const double dt = 0.77 * d->min_dx / (d->c * sqrt(3.0));
    if (!(dt > 0.0) || !(dt < 0.1)) return fail(REQ, "derived dt failed sanity bounds (0, 0.1)");

    AcousticAMRState* amr =
        amr_create(d->n_cells, d->cell_centers, d->cell_volumes, d->open_volume_frac,
                   d->cell_types, d->n_faces, d->face_cell_neg, d->face_cell_pos, d->face_area,
                   d->face_open_frac, d->face_distance, d->c, d->rho_air, d->min_dx,
                   d->gradient_order);
    if (!amr) {
        std::string m = "amr_create returned null";
        const char* stage = amr_get_last_create_error_stage();
        const char* detail = amr_get_last_create_error_message();
        if (stage && stage[0]) m += std::string(" at stage ") + stage;
        if (detail && detail[0]) m += std::string(": ") + detail;
        return fail(AMR_BACKEND_AMR_CREATE_FAILED, m);
    }
    if (amr_setup_plate(amr, d->plate_Nx, d->plate_Ny, d->plate_dx, d->plate_origin,
                        d->plate_active, d->plate_mass_density, d->plate_stiffness_D,
                        d->plate_alpha_M, d->plate_beta_K, d->n_active_plate,
                        d->plate_active_flat_idx, d->plate_face_above_starts,
                        d->plate_face_above_idx, d->plate_face_above_wgt,
                        d->plate_face_below_starts, d->plate_face_below_idx,
                        d->plate_face_below_wgt, d->plate_cell_above,
                        d->plate_cell_below) != SK_OK)
        return fail(AMR_BACKEND_AMR_SETUP_PLATE_FAILED, "amr_setup_plate failed", amr);
    if (amr_set_bridge_plate_sources(amr, d->n_bridge_plate, d->bridge_plate_idx,
                                     d->bridge_plate_wgt) != SK_OK)
        return fail(AMR_BACKEND_BRIDGE_SOURCE_FAILED, "amr_set_bridge_plate_sources failed", amr);
    if (d->n_neck_plate > 0 && d->neck_plate_idx && d->neck_plate_wgt &&
        amr_set_neck_plate_sources(amr, d->n_neck_plate, d->neck_plate_idx,
                                   d->neck_plate_wgt) != SK_OK)
        return fail(AMR_BACKEND_NECK_SOURCE_FAILED, "amr_set_neck_plate_sources failed", amr);
    if (d->n_pml > 0 &&
        amr_set_border_condition(amr, d->border_mode,
                                 d->border_sigma_order > 0.0f ? d->border_sigma_order : 3.0f,
                                 d->border_R_reflection, d->border_Z_match, d->n_pml,
                                 d->bounds_min, d->bounds_max) != SK_OK)
        return fail(AMR_BACKEND_BORDER_SETUP_FAILED, "amr_set_border_condition failed", amr);
    return amr;
}

IPressureBackend* amr_pressure_backend_create(const AMRCoevolverDescriptor* d) {
    AcousticAMRState* amr = amr_state_from_descriptor(d);
    if (!amr) return nullptr;
    auto* b = new (std::nothrow) AMRPressureBackend(amr, d->n_cells, d->plate_Nx, d->plate_Ny);
    if (!b) return fail(AMR_BACKEND_NEW_FAILED, "AMRPressureBackend allocation failed", amr);
    return b;
}
