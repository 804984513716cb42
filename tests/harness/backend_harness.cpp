// backend_harness.cpp — test driver for the pressure-backend interface of
// libafdtd.so (include/afdtd_pressure_backend.hpp).  Reads a descriptor
// written by tests/test_*backend*.py from a directory (meta.txt lists
// "a <name> <bytes>" arrays stored as <name>.bin and "s <name> <value>"
// scalars), builds the AMR backend with amr_pressure_backend_create, then
// per step injects the bridge drive, steps once and samples the mics.
// Writes trace.bin (steps x 4 x n_mics), p.bin, w.bin and uni.bin into the
// same directory.  Exit 3: creation failed (code and message printed).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/afdtd_pressure_backend.hpp"

static std::map<std::string, std::vector<char>> g_arr;
static std::map<std::string, double> g_sc;

template <class T>
static const T* A(const char* k) {
    auto it = g_arr.find(k);
    return it == g_arr.end() ? nullptr : reinterpret_cast<const T*>(it->second.data());
}

static double S(const char* k, double dflt = 0.0) {
    auto it = g_sc.find(k);
    return it == g_sc.end() ? dflt : it->second;
}

static void put(const std::string& path, const void* p, size_t bytes) {
    std::ofstream f(path, std::ios::binary);
    f.write(static_cast<const char*>(p), (std::streamsize)bytes);
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    std::ifstream meta(dir + "/meta.txt");
    std::string kind, name;
    while (meta >> kind >> name) {
        if (kind == "a") {
            size_t bytes;
            meta >> bytes;
            std::vector<char> buf(bytes + 8);
            std::ifstream f(dir + "/" + name + ".bin", std::ios::binary);
            f.read(buf.data(), (std::streamsize)bytes);
            g_arr[name] = std::move(buf);
        } else {
            double v;
            meta >> v;
            g_sc[name] = v;
        }
    }
    AMRCoevolverDescriptor d;
    std::memset(&d, 0, sizeof(d));
    d.n_cells = (int)S("n_cells");
    d.cell_centers = A<double>("cell_centers");
    d.cell_volumes = A<double>("cell_volumes");
    d.open_volume_frac = A<double>("open_volume_frac");
    d.cell_types = A<uint8_t>("cell_types");
    d.cell_levels = A<int16_t>("cell_levels");
    d.n_faces = (int)S("n_faces");
    d.face_cell_neg = A<int32_t>("face_cell_neg");
    d.face_cell_pos = A<int32_t>("face_cell_pos");
    d.face_area = A<double>("face_area");
    d.face_open_frac = A<double>("face_open_frac");
    d.face_distance = A<double>("face_distance");
    d.c = S("c");
    d.rho_air = S("rho_air");
    d.min_dx = S("min_dx");
    for (int k = 0; k < 3; ++k) {
        d.bounds_min[k] = A<double>("bounds")[k];
        d.bounds_max[k] = A<double>("bounds")[3 + k];
    }
    d.plate_Nx = (int)S("plate_Nx");
    d.plate_Ny = (int)S("plate_Ny");
    d.plate_dx = (float)S("plate_dx");
    d.plate_active = A<uint8_t>("plate_active");
    d.plate_mass_density = (float)S("plate_rho_h");
    d.plate_stiffness_D = (float)S("plate_D");
    d.plate_alpha_M = (float)S("plate_alpha_M");
    d.plate_beta_K = (float)S("plate_beta_K");
    d.n_active_plate = (int)S("n_active_plate");
    d.plate_active_flat_idx = A<int32_t>("plate_active_flat_idx");
    d.plate_face_above_starts = A<int32_t>("face_above_starts");
    d.plate_face_above_idx = A<int32_t>("face_above_idx");
    d.plate_face_above_wgt = A<float>("face_above_wgt");
    d.plate_face_below_starts = A<int32_t>("face_below_starts");
    d.plate_face_below_idx = A<int32_t>("face_below_idx");
    d.plate_face_below_wgt = A<float>("face_below_wgt");
    d.plate_cell_above = A<int32_t>("cell_above");
    d.plate_cell_below = A<int32_t>("cell_below");
    d.n_bridge_plate = (int)S("n_bridge");
    d.bridge_plate_idx = A<int32_t>("bridge_idx");
    d.bridge_plate_wgt = A<float>("bridge_wgt");
    d.border_mode = (int)S("border_mode");
    d.border_sigma_order = (float)S("border_sigma_order");
    d.border_R_reflection = (float)S("border_R");
    d.n_pml = (int)S("n_pml");
    d.gradient_order = (int)S("gradient_order", 2);

    IPressureBackend* b = amr_pressure_backend_create(&d);
    if (!b) {
        std::printf("CREATE_FAILED %s %s\n", amr_pressure_backend_last_error_code(),
                    amr_pressure_backend_last_error_message());
        return 3;
    }
    const int n_mics = (int)S("n_mics"), steps = (int)S("steps"), nb = d.n_bridge_plate;
    if (b->setup_mic_samplers(n_mics, A<float>("mic_pos"), A<float>("mic_axis"), 1) != 0) {
        std::printf("MIC_SETUP_FAILED\n");
        return 4;
    }
    std::vector<float> trace((size_t)steps * 4 * n_mics);
    for (int s = 0; s < steps; ++s) {
        if (b->inject_bridge_drive(A<float>("drives") + (size_t)s * nb, nb, (float)S("scale")) != 0)
            return 5;
        const int rc = b->step(1);
        if (rc != 0) {
            std::printf("STEP_FAILED %d at %d\n", rc, s);
            return 6;
        }
        float* t = &trace[(size_t)s * 4 * n_mics];
        if (b->sample_mics(n_mics, t, t + n_mics, t + 2 * n_mics, t + 3 * n_mics) != 0) return 7;
    }
    std::vector<float> p(b->get_pressure_field_size()), w(b->get_plate_displacement_size());
    const int Nx = (int)S("uni_Nx"), Ny = (int)S("uni_Ny"), Nz = (int)S("uni_Nz");
    std::vector<float> uni((size_t)Nx * Ny * Nz);
    int px, py, pc;
    if (b->get_pressure_field_flat(p.data(), (int)p.size()) != 0 ||
        b->get_plate_displacement(w.data(), (int)w.size()) != 0 ||
        b->get_pressure_field_uniform(Nx, Ny, Nz, d.bounds_min, d.bounds_max, uni.data(),
                                      (int)uni.size()) != 0 ||
        b->get_plate_dims(&px, &py, &pc) != 0)
        return 8;
    put(dir + "/trace.bin", trace.data(), trace.size() * 4);
    put(dir + "/p.bin", p.data(), p.size() * 4);
    put(dir + "/w.bin", w.data(), w.size() * 4);
    put(dir + "/uni.bin", uni.data(), uni.size() * 4);
    std::printf("OK dt=%.17g cells=%d plate=%dx%d\n", b->get_dt(), (int)p.size(), px, py);
    delete b;
    return 0;
}
