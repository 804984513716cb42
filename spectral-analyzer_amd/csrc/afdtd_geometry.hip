// afdtd_geometry.hip — the geometry input formats of the uniform solver on the
// device: the body voxeliser (cell types and the plate mask from an XY
// outline polygon, acoustic_fdtd_bridge.py:255-359) and the cut-cell face
// open fractions (acoustic_fdtd_bridge.py:456-552), which feed fdtd_create and
// fdtd_set_face_fractions.
//
// Both are byte work over the grid (one thread per four cells for the cell
// types, one per face for the fractions) plus point-in-polygon tests.  The
// test is the even-odd ray-casting rule evaluated in double with the
// caller's expression order,
//   crossing(edge a -> b) = ((y_a > Y) != (y_b > Y)) &&
//                           X < (x_b - x_a) * (Y - y_a) / (y_b - y_a + 1e-15) + x_a,
// edges taken as (vertex e, vertex e-1), e = 0..n-1 with vertex -1 = n-1, so
// every inside/outside decision equals the host computation bit for bit.
// Every point tested lies on a few lines of constant y (cell-centre rows,
// quarter-cell rows, face rows) and at one of a few x per column, and the
// face sub-points do not depend on z, so the tests are done once per (x, y)
// point into small tables: a workgroup takes one row, collects the x of the
// edges crossing it (xc above) in LDS, and each point's parity is the count
// of crossings with X < xc.
#include <hip/hip_runtime.h>

#include <vector>

#include "afdtd_internal.h"

namespace {

struct GeoGrid {
    int Nx, Ny, Nz;
    double dx, gx, gy, gz;   // cell (i, j, k) has centre g + (index + 0.5) dx
};

constexpr int MAX_PTS = 4096;   // outline vertices (crossings of one row kept in LDS)

// Point-in-polygon tables.  TAB 0: cell centres, out[i*Ny + j];
// TAB 1 (Vz face sub-points): x = i + 0.25|0.75, y = j + 0.25|0.75,
//   out[(2i+a)*(2Ny) + 2j+b];
// TAB 2 (Vx faces): x = i + 1 (i < Nx-1), y = j + 0.25|0.75, out[i*(2Ny) + 2j+a];
// TAB 3 (Vy faces): x = i + 0.25|0.75, y = j + 1 (j < Ny-1), out[(2i+a)*(Ny-1) + j]
// (positions in cells from the grid origin).  Block = one row of constant y.
template <int TAB>
__global__ void __launch_bounds__(256)
k_pip_table(GeoGrid g, const float2* __restrict__ poly, int n, uint8_t* __restrict__ out) {
    __shared__ double xc[MAX_PTS];
    __shared__ int ncross;
    const int r = blockIdx.x, tid = threadIdx.x;
    const double qo[2] = {0.25, 0.75};
    double Y;
    int nrows, ncols;
    if (TAB == 0) {
        Y = g.gy + ((double)r + 0.5) * g.dx;
        nrows = g.Ny; ncols = g.Nx;
    } else if (TAB == 1 || TAB == 2) {
        Y = g.gy + ((double)(r >> 1) + qo[r & 1]) * g.dx;
        nrows = 2 * g.Ny; ncols = TAB == 1 ? 2 * g.Nx : g.Nx - 1;
    } else {
        Y = g.gy + (double)(r + 1) * g.dx;
        nrows = g.Ny - 1; ncols = 2 * g.Nx;
    }
    if (tid == 0) ncross = 0;
    __syncthreads();
    for (int e = tid; e < n; e += 256) {
        const double xi = (double)poly[e].x, yi = (double)poly[e].y;
        const int pe = e == 0 ? n - 1 : e - 1;
        const double xj = (double)poly[pe].x, yj = (double)poly[pe].y;
        if ((yi > Y) != (yj > Y)) xc[atomicAdd(&ncross, 1)] = (xj - xi) * (Y - yi) / (yj - yi + 1e-15) + xi;
    }
    __syncthreads();
    const int m = ncross;
    for (int c = tid; c < ncols; c += 256) {
        double X;
        if (TAB == 0) X = g.gx + ((double)c + 0.5) * g.dx;
        else if (TAB == 2) X = g.gx + (double)(c + 1) * g.dx;
        else X = g.gx + ((double)(c >> 1) + qo[c & 1]) * g.dx;
        bool inside = false;
        for (int q = 0; q < m; ++q) inside ^= X < xc[q];
        out[(size_t)c * nrows + r] = inside ? 1 : 0;
    }
}

// Cell types, four consecutive cells per thread: inside columns get the back
// wall at iz_back and the plate at iz_top (the plate wins when they
// coincide); outside columns are walls wherever z_min - dx < z < z_max + dx.
__global__ void __launch_bounds__(256)
k_voxel_fill(GeoGrid g, const uint8_t* __restrict__ inside, double zlo, double zhi, int iz_back,
             int iz_top, uint8_t* __restrict__ cell_type) {
    // grids hold fewer than 2^31 cells, so 32-bit index arithmetic suffices
    const uint32_t N = (uint32_t)g.Nx * g.Ny * g.Nz;
    const uint32_t c0 = (blockIdx.x * 256u + threadIdx.x) * 4u;
    if (c0 >= N) return;
    uint32_t word = 0;
    const int nv = (int)(N - c0 < 4 ? N - c0 : 4);
    uint32_t col = c0 / (uint32_t)g.Nz;
    int k = (int)(c0 - col * (uint32_t)g.Nz);
    for (int u = 0; u < nv; ++u, ++k) {
        if (k == g.Nz) { k = 0; ++col; }
        uint8_t t = FDTD_AIR;
        if (inside[col]) {
            if (k == iz_top) t = FDTD_PLATE;
            else if (k == iz_back) t = FDTD_WALL;
        } else {
            const double z = g.gz + ((double)k + 0.5) * g.dx;
            if (zlo < z && z < zhi) t = FDTD_WALL;
        }
        word |= (uint32_t)t << (8 * u);
    }
    if (nv == 4 && ((uintptr_t)(cell_type + c0) & 3) == 0) {
        *(uint32_t*)(cell_type + c0) = word;
    } else {
        for (int u = 0; u < nv; ++u) cell_type[c0 + u] = (uint8_t)(word >> (8 * u));
    }
}

__device__ __forceinline__ bool solid(uint8_t t) { return t == FDTD_WALL || t == FDTD_PLATE; }

// Open fraction of every face normal to axis AX (0 = x, 1 = y, 2 = z), in the
// packed staggered layout of that component.  A face between two solid or
// two open cells is 0 or 1; a face with exactly one solid neighbour is the
// share of its four sub-points (offsets 0.25 and 0.75 of a cell along the two
// tangential axes, first tangential axis outer) that lie inside the outline
// (table `tab` of k_pip_table) and within [zb_lo, zb_hi] in z.  Vz faces of
// the two layers next to the plate (plate_iz - 1 and plate_iz, when
// 0 < plate_iz < Nz - 1) are 1 under an active plate node and keep the
// solid/open value elsewhere.
template <int AX>
__global__ void __launch_bounds__(256)
k_face_fractions(GeoGrid g, const uint8_t* __restrict__ ct, const uint8_t* __restrict__ pa,
                 const uint8_t* __restrict__ tab, double zb_lo, double zb_hi, int plate_iz,
                 float* __restrict__ frac) {
    const uint32_t Fx = g.Nx - (AX == 0), Fy = g.Ny - (AX == 1), Fz = g.Nz - (AX == 2);
    const uint32_t F = Fx * Fy * Fz;                  // < 2^31
    const uint32_t f = blockIdx.x * 256u + threadIdx.x;
    if (f >= F) return;
    const uint32_t t = f / Fz;
    const int k = (int)(f - t * Fz);
    const uint32_t i = t / Fy;
    const int j = (int)(t - i * Fy);
    const uint32_t sy = g.Nz, sx = (uint32_t)g.Ny * g.Nz;
    const uint32_t cL = (uint32_t)k + sy * j + sx * i;
    const uint32_t cR = cL + (AX == 0 ? sx : AX == 1 ? sy : 1u);
    const bool sL = solid(ct[cL]), sR = solid(ct[cR]);
    float v = (sL || sR) ? 0.0f : 1.0f;
    bool cut = sL != sR;
    if (AX == 2 && plate_iz > 0 && plate_iz < g.Nz - 1 && (k == plate_iz - 1 || k == plate_iz)) {
        cut = false;
        if (pa[(size_t)i * g.Ny + j]) v = 1.0f;
    }
    if (cut) {
        const double o[2] = {0.25, 0.75};
        float s = 0.0f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                double Z;
                bool in_xy;
                if (AX == 0) {
                    Z = g.gz + ((double)k + o[b]) * g.dx;
                    in_xy = tab[(size_t)i * (2 * g.Ny) + 2 * j + a];
                } else if (AX == 1) {
                    Z = g.gz + ((double)k + o[b]) * g.dx;
                    in_xy = tab[(size_t)(2 * i + a) * (g.Ny - 1) + j];
                } else {
                    Z = g.gz + (double)(k + 1) * g.dx;
                    in_xy = tab[(size_t)(2 * i + a) * (2 * g.Ny) + 2 * j + b];
                }
                s += (Z >= zb_lo && Z <= zb_hi && in_xy) ? 1.0f : 0.0f;
            }
        v = s / 4.0f;
    }
    frac[f] = v;
}

int blocks_for(size_t n, size_t per_block) { return (int)((n + per_block - 1) / per_block); }

// Kernels of one entry point on `s`, bracketed by events when the caller
// asks for their time; the call returns when they are done.
struct Timed {
    hipStream_t s;
    float* ms;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Timed(hipStream_t s_, float* ms_) : s(s_), ms(ms_) {
        if (ms && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
                   hipEventRecord(e0, s) != hipSuccess))
            ms = nullptr;
    }
    hipError_t finish() {
        if (ms && hipEventRecord(e1, s) != hipSuccess) ms = nullptr;
        hipError_t e = hipStreamSynchronize(s);
        if (e == hipSuccess && ms) e = hipEventElapsedTime(ms, e0, e1);
        return e;
    }
    ~Timed() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

bool geo_ok(int Nx, int Ny, int Nz, double dx, int n_pts) {
    return Nx >= 2 && Ny >= 2 && Nz >= 2 && dx > 0.0 && n_pts >= 3 && n_pts <= MAX_PTS &&
           (double)Nx * Ny * Nz < 2147483647.0;
}

}  // namespace

extern "C" {

SK_API int fdtd_voxelise_body_device(const float* outline_xy, int n_pts, int Nx, int Ny, int Nz,
                                     double gx_min, double gy_min, double gz_min, double dx,
                                     double z_min, double z_max, int iz_back, int iz_top,
                                     unsigned char* cell_type, unsigned char* plate_active,
                                     void* stream, float* kernel_ms) {
    if (!outline_xy || !cell_type || !plate_active) return FDTD_ERR_NULL;
    if (!geo_ok(Nx, Ny, Nz, dx, n_pts)) return FDTD_ERR_DIM;
    const GeoGrid g{Nx, Ny, Nz, dx, gx_min, gy_min, gz_min};
    hipStream_t s = (hipStream_t)stream;
    Timed tm(s, kernel_ms);
    const size_t cols = (size_t)Nx * Ny, N = cols * Nz;
    hipLaunchKernelGGL(k_pip_table<0>, dim3(Ny), dim3(256), 0, s, g, (const float2*)outline_xy,
                       n_pts, plate_active);
    hipLaunchKernelGGL(k_voxel_fill, dim3(blocks_for(N, 1024)), dim3(256), 0, s, g,
                       (const uint8_t*)plate_active, z_min - dx, z_max + dx, iz_back, iz_top,
                       cell_type);
    AFDTD_CHECK(hipGetLastError());
    AFDTD_CHECK(tm.finish());
    return FDTD_OK;
}

SK_API int fdtd_cut_cell_fractions_device(const unsigned char* cell_type,
                                          const unsigned char* plate_active,
                                          const float* outline_xy, int n_pts, int Nx, int Ny,
                                          int Nz, double gx_min, double gy_min, double gz_min,
                                          double dx, int plate_iz, int iz_back, float* vx_frac,
                                          float* vy_frac, float* vz_frac, void* stream,
                                          float* kernel_ms) {
    if (!cell_type || !plate_active || !outline_xy || !vx_frac || !vy_frac || !vz_frac)
        return FDTD_ERR_NULL;
    if (!geo_ok(Nx, Ny, Nz, dx, n_pts)) return FDTD_ERR_DIM;
    const GeoGrid g{Nx, Ny, Nz, dx, gx_min, gy_min, gz_min};
    hipStream_t s = (hipStream_t)stream;
    Timed tm(s, kernel_ms);
    const double zb_lo = gz_min + (double)iz_back * dx;
    const double zb_hi = gz_min + ((double)plate_iz + 0.5) * dx;
    const float2* poly = (const float2*)outline_xy;
    const size_t Fx = (size_t)(Nx - 1) * Ny * Nz, Fy = (size_t)Nx * (Ny - 1) * Nz,
                 Fz = (size_t)Nx * Ny * (Nz - 1);
    // sub-point tables of the three face families
    const size_t nT[3] = {(size_t)(Nx - 1) * 2 * Ny, (size_t)2 * Nx * (Ny - 1),
                          (size_t)4 * Nx * Ny};
    uint8_t* tab = nullptr;
    AFDTD_CHECK(hipMalloc(&tab, nT[0] + nT[1] + nT[2]));
    uint8_t *tx = tab, *ty = tab + nT[0], *tz = ty + nT[1];
    hipLaunchKernelGGL(k_pip_table<2>, dim3(2 * Ny), dim3(256), 0, s, g, poly, n_pts, tx);
    hipLaunchKernelGGL(k_pip_table<3>, dim3(Ny - 1), dim3(256), 0, s, g, poly, n_pts, ty);
    hipLaunchKernelGGL(k_pip_table<1>, dim3(2 * Ny), dim3(256), 0, s, g, poly, n_pts, tz);
    hipLaunchKernelGGL(k_face_fractions<0>, dim3(blocks_for(Fx, 256)), dim3(256), 0, s, g,
                       cell_type, plate_active, tx, zb_lo, zb_hi, plate_iz, vx_frac);
    hipLaunchKernelGGL(k_face_fractions<1>, dim3(blocks_for(Fy, 256)), dim3(256), 0, s, g,
                       cell_type, plate_active, ty, zb_lo, zb_hi, plate_iz, vy_frac);
    hipLaunchKernelGGL(k_face_fractions<2>, dim3(blocks_for(Fz, 256)), dim3(256), 0, s, g,
                       cell_type, plate_active, tz, zb_lo, zb_hi, plate_iz, vz_frac);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = tm.finish();
    (void)hipFree(tab);
    AFDTD_CHECK(e);
    return FDTD_OK;
}

// Host-memory forms: copy in, run on the device, copy out (synchronous).
SK_API int fdtd_voxelise_body(const float* outline_xy, int n_pts, int Nx, int Ny, int Nz,
                              double gx_min, double gy_min, double gz_min, double dx,
                              double z_min, double z_max, int iz_back, int iz_top,
                              unsigned char* cell_type, unsigned char* plate_active) {
    if (!outline_xy || !cell_type || !plate_active) return FDTD_ERR_NULL;
    if (!geo_ok(Nx, Ny, Nz, dx, n_pts)) return FDTD_ERR_DIM;
    const size_t cols = (size_t)Nx * Ny, N = cols * Nz, pb = sizeof(float) * 2 * n_pts;
    void *dpoly = nullptr, *dct = nullptr, *dpa = nullptr;
    int rc = FDTD_ERR_DEVICE;
    if (hipMalloc(&dpoly, pb) == hipSuccess && hipMalloc(&dct, N) == hipSuccess &&
        hipMalloc(&dpa, cols) == hipSuccess &&
        hipMemcpy(dpoly, outline_xy, pb, hipMemcpyHostToDevice) == hipSuccess) {
        rc = fdtd_voxelise_body_device((const float*)dpoly, n_pts, Nx, Ny, Nz, gx_min, gy_min,
                                       gz_min, dx, z_min, z_max, iz_back, iz_top,
                                       (unsigned char*)dct, (unsigned char*)dpa, nullptr,
                                       nullptr);
        if (rc == FDTD_OK &&
            (hipMemcpy(cell_type, dct, N, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(plate_active, dpa, cols, hipMemcpyDeviceToHost) != hipSuccess))
            rc = FDTD_ERR_DEVICE;
    }
    if (rc == FDTD_ERR_DEVICE) afdtd_set_error("fdtd_voxelise_body: device allocation or copy failed");
    (void)hipFree(dpoly);
    (void)hipFree(dct);
    (void)hipFree(dpa);
    return rc;
}

SK_API int fdtd_cut_cell_fractions(const unsigned char* cell_type,
                                   const unsigned char* plate_active, const float* outline_xy,
                                   int n_pts, int Nx, int Ny, int Nz, double gx_min,
                                   double gy_min, double gz_min, double dx, int plate_iz,
                                   int iz_back, float* vx_frac, float* vy_frac, float* vz_frac) {
    if (!cell_type || !plate_active || !outline_xy || !vx_frac || !vy_frac || !vz_frac)
        return FDTD_ERR_NULL;
    if (!geo_ok(Nx, Ny, Nz, dx, n_pts)) return FDTD_ERR_DIM;
    const size_t cols = (size_t)Nx * Ny, N = cols * Nz, pb = sizeof(float) * 2 * n_pts;
    const size_t nf[3] = {(size_t)(Nx - 1) * Ny * Nz, (size_t)Nx * (Ny - 1) * Nz,
                          (size_t)Nx * Ny * (Nz - 1)};
    void *dpoly = nullptr, *dct = nullptr, *dpa = nullptr, *df[3] = {nullptr, nullptr, nullptr};
    bool ok = hipMalloc(&dpoly, pb) == hipSuccess && hipMalloc(&dct, N) == hipSuccess &&
              hipMalloc(&dpa, cols) == hipSuccess;
    for (int a = 0; a < 3; ++a) ok = ok && hipMalloc(&df[a], sizeof(float) * nf[a]) == hipSuccess;
    ok = ok && hipMemcpy(dpoly, outline_xy, pb, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dct, cell_type, N, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(dpa, plate_active, cols, hipMemcpyHostToDevice) == hipSuccess;
    int rc = FDTD_ERR_DEVICE;
    if (ok) {
        rc = fdtd_cut_cell_fractions_device((const unsigned char*)dct, (const unsigned char*)dpa,
                                            (const float*)dpoly, n_pts, Nx, Ny, Nz, gx_min, gy_min,
                                            gz_min, dx, plate_iz, iz_back, (float*)df[0],
                                            (float*)df[1], (float*)df[2], nullptr, nullptr);
        float* out[3] = {vx_frac, vy_frac, vz_frac};
        for (int a = 0; a < 3 && rc == FDTD_OK; ++a)
            if (hipMemcpy(out[a], df[a], sizeof(float) * nf[a], hipMemcpyDeviceToHost) != hipSuccess)
                rc = FDTD_ERR_DEVICE;
    }
    if (rc == FDTD_ERR_DEVICE)
        afdtd_set_error("fdtd_cut_cell_fractions: device allocation or copy failed");
    (void)hipFree(dpoly);
    (void)hipFree(dct);
    (void)hipFree(dpa);
    for (int a = 0; a < 3; ++a) (void)hipFree(df[a]);
    return rc;
}

}  // extern "C"
