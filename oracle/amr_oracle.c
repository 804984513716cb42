/*
 * amr_oracle.c — CPU oracle for the AMR (topology-driven, CSR) acoustic
 * stepper, gradient order 2.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/ and the cpu_baseline leg of bench.py
 * load this library, as the checker.  The product (libafdtd.so) never links,
 * calls or falls back to it.
 *
 * A plain-C restatement of the reference's AMR stepper
 * (csrc/kernels/acoustic_amr.cpp): set-up precompute (:700-855), the
 * order-2 step (:1100-1204), the border condition (:1318-1409), nearest-cell
 * pressure injection (:1227-1244), field getters (:1246-1302) and the
 * uniform-grid scatter (:1266-1294).  Float arithmetic follows the
 * reference's expression order and is compiled without FMA contraction, so
 * the HIP kernels are expected to equal it bit for bit.
 *
 * Not restated: gradient order 8 (the Fornberg stencil builder,
 * :398-631), the AMR plate (:956-1048) and the mic samplers (:1647-1792);
 * amo_create returns NULL for order 8.
 *
 * The reference cannot be compiled here (acoustic_amr.cpp includes
 * <Eigen/Dense>, which the image lacks), so this restatement is pinned by a
 * hand-derived one-step answer, exact flux conservation, reduction to the
 * uniform Yee oracle on a box grid and the border condition's special cases
 * (tests/test_amr_oracle_cpu.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SK_EXPORT __attribute__((visibility("default")))
#define AMO_OK 0
#define AMO_ERR_NULL (-1)
#define AMO_ERR_DIM (-2)
#define AMO_ERR_DIVERGED (-3)

typedef struct Amo {
    int nc, nf, order, step_count;
    double c, rho, dt;
    double *centers;          /* nc*3 */
    double *open_vol;         /* nc */
    uint8_t *type;            /* nc */
    int *neg, *pos;           /* nf */
    float *p, *v;             /* fields */
    float *inv_d, *ffc;       /* per face (:800-804) */
    float *inv_den;           /* per cell (:806-811) */
    float *pdamp, *psrc, *vdamp, *alpha;
    int *acoustic; int n_acoustic;   /* (:783-785) */
    int *wallplate; int n_wallplate; /* (:813-816) */
    int *csr_start, *csr_face; float *csr_w;  /* (:818-849) */
} Amo;

/* Input validation, acoustic_amr.cpp:275-311. */
static int amo_validate(int nc, const double* cc, const double* vol, const double* ov,
                        const uint8_t* ct, int nf, const int32_t* fn, const int32_t* fp,
                        const double* fa, const double* fo, const double* fd,
                        double c, double rho, double min_dx)
{
    if (nc <= 0 || nf <= 0) return AMO_ERR_DIM;
    if (!cc || !vol || !ov || !ct || !fn || !fp || !fa || !fo || !fd) return AMO_ERR_NULL;
    if (c <= 0.0 || rho <= 0.0 || min_dx <= 0.0) return AMO_ERR_DIM;
    for (int i = 0; i < nc; ++i) {
        if (!(vol[i] > 0.0)) return AMO_ERR_DIM;
        if (ov[i] < 0.0 || ov[i] > 1.0) return AMO_ERR_DIM;
        if (ct[i] > 3) return AMO_ERR_DIM;
    }
    for (int f = 0; f < nf; ++f) {
        if (fn[f] < 0 || fn[f] >= nc || fp[f] < 0 || fp[f] >= nc) return AMO_ERR_DIM;
        if (fn[f] == fp[f]) return AMO_ERR_DIM;
        if (!(fa[f] > 0.0)) return AMO_ERR_DIM;
        if (fo[f] < 0.0 || fo[f] > 1.0) return AMO_ERR_DIM;
        if (!(fd[f] > 0.0)) return AMO_ERR_DIM;
    }
    return AMO_OK;
}

SK_EXPORT void amo_destroy(Amo* a)
{
    if (!a) return;
    free(a->centers); free(a->open_vol); free(a->type); free(a->neg); free(a->pos);
    free(a->p); free(a->v); free(a->inv_d); free(a->ffc); free(a->inv_den);
    free(a->pdamp); free(a->psrc); free(a->vdamp); free(a->alpha);
    free(a->acoustic); free(a->wallplate);
    free(a->csr_start); free(a->csr_face); free(a->csr_w);
    free(a);
}

/* amr_create, acoustic_amr.cpp:682-880 (order 2 only). */
SK_EXPORT Amo* amo_create(int nc, const double* cc, const double* vol, const double* ov,
                          const uint8_t* ct, int nf, const int32_t* fn, const int32_t* fp,
                          const double* fa, const double* fo, const double* fd,
                          double c, double rho, double min_dx, int order)
{
    if (amo_validate(nc, cc, vol, ov, ct, nf, fn, fp, fa, fo, fd, c, rho, min_dx) != AMO_OK)
        return NULL;
    if (order != 2) return NULL;
    Amo* a = (Amo*)calloc(1, sizeof(Amo));
    a->nc = nc; a->nf = nf; a->order = order;
    a->c = c; a->rho = rho;
    a->dt = 0.77 * min_dx / (c * sqrt(3.0));                         /* :744-745 */
    a->centers = (double*)malloc(sizeof(double) * 3 * (size_t)nc);
    memcpy(a->centers, cc, sizeof(double) * 3 * (size_t)nc);
    a->open_vol = (double*)malloc(sizeof(double) * (size_t)nc);
    memcpy(a->open_vol, ov, sizeof(double) * (size_t)nc);
    a->type = (uint8_t*)malloc((size_t)nc);
    memcpy(a->type, ct, (size_t)nc);
    a->neg = (int*)malloc(sizeof(int) * (size_t)nf);
    a->pos = (int*)malloc(sizeof(int) * (size_t)nf);
    memcpy(a->neg, fn, sizeof(int) * (size_t)nf);
    memcpy(a->pos, fp, sizeof(int) * (size_t)nf);
    a->p = (float*)calloc((size_t)nc, sizeof(float));
    a->v = (float*)calloc((size_t)nf, sizeof(float));
    a->acoustic = (int*)malloc(sizeof(int) * (size_t)nc);
    a->wallplate = (int*)malloc(sizeof(int) * (size_t)nc);
    for (int i = 0; i < nc; ++i)                                      /* :783-785 */
        if ((ct[i] == 0 || ct[i] == 3) && ov[i] > 0.0) a->acoustic[a->n_acoustic++] = i;
    if (a->n_acoustic == 0) { amo_destroy(a); return NULL; }          /* :794-798 */
    a->inv_d = (float*)malloc(sizeof(float) * (size_t)nf);
    a->ffc = (float*)malloc(sizeof(float) * (size_t)nf);
    for (int f = 0; f < nf; ++f) {                                    /* :801-804 */
        a->inv_d[f] = 1.0f / (float)fd[f];
        a->ffc[f] = (float)(fa[f] * fo[f]);
    }
    a->inv_den = (float*)calloc((size_t)nc, sizeof(float));
    for (int t = 0; t < a->n_acoustic; ++t) {                        /* :808-811 */
        int idx = a->acoustic[t];
        double den = vol[idx] * fmax(ov[idx], 1e-12);
        a->inv_den[idx] = (float)(1.0 / den);
    }
    for (int i = 0; i < nc; ++i)                                      /* :813-816 */
        if (ct[i] == 1 || ct[i] == 2) a->wallplate[a->n_wallplate++] = i;
    a->csr_start = (int*)calloc((size_t)nc + 1, sizeof(int));          /* :818-831 */
    for (int f = 0; f < nf; ++f) { a->csr_start[fn[f] + 1]++; a->csr_start[fp[f] + 1]++; }
    for (int i = 1; i <= nc; ++i) a->csr_start[i] += a->csr_start[i - 1];
    int ne = a->csr_start[nc];
    a->csr_face = (int*)malloc(sizeof(int) * (size_t)ne);
    a->csr_w = (float*)malloc(sizeof(float) * (size_t)ne);
    int* fill = (int*)calloc((size_t)nc, sizeof(int));
    for (int f = 0; f < nf; ++f) {                                    /* :837-848 */
        float w = a->ffc[f];
        int ka = a->csr_start[fn[f]] + fill[fn[f]]++;
        a->csr_face[ka] = f; a->csr_w[ka] = +w;
        int kb = a->csr_start[fp[f]] + fill[fp[f]]++;
        a->csr_face[kb] = f; a->csr_w[kb] = -w;
    }
    free(fill);
    a->alpha = (float*)calloc((size_t)nc, sizeof(float));             /* :852-859 */
    a->vdamp = (float*)malloc(sizeof(float) * (size_t)nf);
    a->pdamp = (float*)malloc(sizeof(float) * (size_t)nc);
    a->psrc = (float*)malloc(sizeof(float) * (size_t)nc);
    for (int f = 0; f < nf; ++f) a->vdamp[f] = 1.0f;
    for (int i = 0; i < nc; ++i) { a->pdamp[i] = 1.0f; a->psrc[i] = 1.0f; }
    return a;
}

SK_EXPORT double amo_get_dt(const Amo* a) { return a ? a->dt : 0.0; }
SK_EXPORT int amo_get_step_count(const Amo* a) { return a ? a->step_count : 0; }
SK_EXPORT int amo_csr_entries(const Amo* a) { return a ? a->csr_start[a->nc] : 0; }

/* amr_reset, acoustic_amr.cpp:924-936 (no plate here). */
SK_EXPORT int amo_reset(Amo* a)
{
    if (!a) return AMO_ERR_NULL;
    memset(a->p, 0, sizeof(float) * (size_t)a->nc);
    memset(a->v, 0, sizeof(float) * (size_t)a->nf);
    a->step_count = 0;
    return AMO_OK;
}

/* Test hook: load a state (no reference counterpart). */
SK_EXPORT int amo_set_state(Amo* a, const float* p, const float* v)
{
    if (!a) return AMO_ERR_NULL;
    if (p) memcpy(a->p, p, sizeof(float) * (size_t)a->nc);
    if (v) memcpy(a->v, v, sizeof(float) * (size_t)a->nf);
    return AMO_OK;
}

/* amr_step, order-2 branch, acoustic_amr.cpp:1100-1204 (plate hooks are
 * no-ops without a plate). */
SK_EXPORT int amo_step(Amo* a, int n_steps)
{
    if (!a) return AMO_ERR_NULL;
    if (n_steps < 0) return AMO_ERR_DIM;
    if (n_steps == 0) return AMO_OK;
    const float dt = (float)a->dt;                                    /* :1111-1114 */
    const float rho = (float)a->rho;
    const float bulk = (float)(a->rho * a->c * a->c * a->dt);
    const float dt_over_rho = dt / rho;
    const int nf = a->nf, nc = a->nc;
    for (int s = 0; s < n_steps; ++s) {
        #pragma omp parallel for schedule(static)
        for (int f = 0; f < nf; ++f) {                                /* :1147-1154 */
            const float g = (a->p[a->pos[f]] - a->p[a->neg[f]]) * a->inv_d[f];
            a->v[f] = (a->v[f] - dt_over_rho * g) * a->vdamp[f];
        }
        #pragma omp parallel for schedule(static)
        for (int c = 0; c < nc; ++c) {                                /* :1163-1172 */
            float d = 0.0f;
            for (int k = a->csr_start[c]; k < a->csr_start[c + 1]; ++k)
                d += a->csr_w[k] * a->v[a->csr_face[k]];
            a->p[c] = a->p[c] * a->pdamp[c] - bulk * d * a->inv_den[c] * a->psrc[c];
        }
        for (int t = 0; t < a->n_wallplate; ++t) a->p[a->wallplate[t]] = 0.0f;  /* :1174-1176 */
        int bad = 0;                                                  /* :1181-1192 */
        for (int c = 0; c < nc; ++c) if (!isfinite(a->p[c])) { bad = 1; break; }
        if (!bad && (a->step_count & 0xFF) == 0) {
            float pmax = 0.0f;
            for (int c = 0; c < nc; ++c) { float q = fabsf(a->p[c]); if (q > pmax) pmax = q; }
            if (pmax > 2e3f) bad = 1;
        }
        if (bad) return AMO_ERR_DIVERGED;
        ++a->step_count;
    }
    return AMO_OK;
}

/* amr_inject_pressure_nearest, acoustic_amr.cpp:1227-1244. */
SK_EXPORT int amo_inject_nearest(Amo* a, const double* xyz, float value)
{
    if (!a || !xyz) return AMO_ERR_NULL;
    if (!isfinite(value)) return AMO_ERR_DIVERGED;
    double best = 1e300;
    int best_idx = -1;
    for (int t = 0; t < a->n_acoustic; ++t) {
        int idx = a->acoustic[t];
        double dx = a->centers[3 * idx + 0] - xyz[0];
        double dy = a->centers[3 * idx + 1] - xyz[1];
        double dz = a->centers[3 * idx + 2] - xyz[2];
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best) { best = d2; best_idx = idx; }
    }
    if (best_idx < 0) return AMO_ERR_DIM;
    a->p[best_idx] += value;
    return AMO_OK;
}

/* amr_get_pressure / amr_get_velocity, acoustic_amr.cpp:1246-1252, :1296-1302. */
SK_EXPORT int amo_get_pressure(const Amo* a, float* out, int n)
{
    if (!a || !out) return AMO_ERR_NULL;
    if (n != a->nc) return AMO_ERR_DIM;
    memcpy(out, a->p, sizeof(float) * (size_t)n);
    return AMO_OK;
}

SK_EXPORT int amo_get_velocity(const Amo* a, float* out, int n)
{
    if (!a || !out) return AMO_ERR_NULL;
    if (n != a->nf) return AMO_ERR_DIM;
    memcpy(out, a->v, sizeof(float) * (size_t)n);
    return AMO_OK;
}

/* Coefficient dump for the pins (no reference counterpart). */
SK_EXPORT int amo_get_coeffs(const Amo* a, float* pdamp, float* psrc, float* vdamp)
{
    if (!a) return AMO_ERR_NULL;
    if (pdamp) memcpy(pdamp, a->pdamp, sizeof(float) * (size_t)a->nc);
    if (psrc) memcpy(psrc, a->psrc, sizeof(float) * (size_t)a->nc);
    if (vdamp) memcpy(vdamp, a->vdamp, sizeof(float) * (size_t)a->nf);
    return AMO_OK;
}

/* amr_scatter_pressure_uniform, acoustic_amr.cpp:1266-1294. */
SK_EXPORT int amo_scatter_uniform(const Amo* a, int Nx, int Ny, int Nz,
                                  const double* bmin, const double* bmax, float* out)
{
    if (!a || !out || !bmin || !bmax) return AMO_ERR_NULL;
    if (Nx <= 0 || Ny <= 0 || Nz <= 0) return AMO_ERR_DIM;
    memset(out, 0, sizeof(float) * (size_t)Nx * Ny * Nz);
    const double rx = (Nx > 1) ? (Nx - 1) / (bmax[0] - bmin[0]) : 0.0;
    const double ry = (Ny > 1) ? (Ny - 1) / (bmax[1] - bmin[1]) : 0.0;
    const double rz = (Nz > 1) ? (Nz - 1) / (bmax[2] - bmin[2]) : 0.0;
    for (int ci = 0; ci < a->nc; ++ci) {
        int ix = (int)lround((a->centers[3 * ci + 0] - bmin[0]) * rx);
        int iy = (int)lround((a->centers[3 * ci + 1] - bmin[1]) * ry);
        int iz = (int)lround((a->centers[3 * ci + 2] - bmin[2]) * rz);
        if (ix < 0 || ix >= Nx || iy < 0 || iy >= Ny || iz < 0 || iz >= Nz) continue;
        out[(size_t)ix * Ny * Nz + (size_t)iy * Nz + iz] = a->p[ci];
    }
    return AMO_OK;
}

/* amr_set_border_condition, acoustic_amr.cpp:1318-1409. */
SK_EXPORT int amo_set_border_condition(Amo* a, int mode, float sigma_order, float R,
                                       float z_match, int n_pml,
                                       const double* bmin, const double* bmax)
{
    (void)z_match;
    if (!a || !bmin || !bmax) return AMO_ERR_NULL;
    if (n_pml <= 0) return AMO_OK;
    const float eff_order = (sigma_order > 0.0f) ? sigma_order : 3.0f;
    const float dt_f = (float)a->dt;
    float sigma_max = 3.45f / (dt_f * (float)n_pml);
    if (mode == 1) {
        float r = fmaxf(0.0f, fminf(1.0f, R));
        sigma_max *= (1.0f - r);
    }
    const double min_dx = a->dt * a->c * sqrt(3.0) / 0.77;
    const double depth = (double)n_pml * min_dx;
    for (int i = 0; i < a->nc; ++i) {
        a->alpha[i] = 0.0f;
        if (a->type[i] != 3) continue;
        const double cx = a->centers[3 * i], cy = a->centers[3 * i + 1], cz = a->centers[3 * i + 2];
        double d = cx - bmin[0];
        if (bmax[0] - cx < d) d = bmax[0] - cx;
        if (cy - bmin[1] < d) d = cy - bmin[1];
        if (bmax[1] - cy < d) d = bmax[1] - cy;
        if (cz - bmin[2] < d) d = cz - bmin[2];
        if (bmax[2] - cz < d) d = bmax[2] - cz;
        if (d < 0.0) d = 0.0;
        double t = 1.0 - d / depth;
        if (t > 1.0) t = 1.0;
        if (t < 0.0) t = 0.0;
        const float norm = (float)t;
        a->alpha[i] = sigma_max * powf(norm, eff_order);
    }
    for (int i = 0; i < a->nc; ++i) {
        const float s = a->alpha[i] * dt_f;
        a->pdamp[i] = expf(-s);
        a->psrc[i] = (s < 1e-6f) ? 1.0f - 0.5f * s : (1.0f - expf(-s)) / s;
    }
    for (int f = 0; f < a->nf; ++f) {
        const int ni = a->neg[f], pi = a->pos[f];
        const float vn = (float)a->open_vol[ni];
        const float vp = (float)a->open_vol[pi];
        const float den = vn + vp;
        const float sf = (den > 0.f)
            ? 2.0f * (vp * a->alpha[ni] + vn * a->alpha[pi]) / den
            : 0.5f * (a->alpha[ni] + a->alpha[pi]);
        a->vdamp[f] = expf(-sf * dt_f);
    }
    return AMO_OK;
}
