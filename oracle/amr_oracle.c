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
 * Gradient order 8 (the Fornberg stencil builder, :313-679) is restated
 * too; its progress reporting is not.
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
    int *sc; float *sw;                       /* order 8: 8 cells + weights per face */
    int *dnb[6];                              /* order 8: +x,-x,+y,-y,+z,-z neighbours */
    double *volumes;                          /* nc (mic search radius) */
    double *face_area, *face_open;            /* nf (mic velocity weights) */
    /* Kirchhoff plate (:1412-1531) */
    int plate_on, pNx, pNy, n_pnodes, n_active;
    float p_dx, p_rh, p_D, p_aM, p_bK;
    float *w, *w_prev, *w_new, *ext;
    int *active;                              /* flat plate index per active node */
    int *ab_start, *ab_face, *ab_side;        /* faces above: CSR over nodes, air cell */
    float *ab_wgt;
    int *bl_start, *bl_face, *bl_side;        /* faces below */
    float *bl_wgt;
    int *cell_above, *cell_below;
    int n_bridge, n_neck;
    int *bridge_idx, *neck_idx;
    float *bridge_wgt, *neck_wgt;
    /* mic samplers (:1647-1760): per mic growing lists, as the reference's
     * vectors (a second set-up appends) */
    int n_mics;
    int *mp_n, **mp_cell; float **mp_w;       /* pressure: cells, weights */
    int *mv_n, **mv_face; float **mv_wx, **mv_wy, **mv_wz;
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

/* ── Gradient order 8 (acoustic_amr.cpp:313-679) ───────────────────────── */

/* Fornberg's first-derivative weights at xi over N <= 8 nodes, float, the
 * reference's loop (acoustic_amr.cpp:398-426). */
SK_EXPORT void amo_fornberg_d1(const float* x, int N, float xi, float* w)
{
    float c[16];
    for (int k = 0; k < 16; ++k) c[k] = 0.0f;
#define CC(k, m) c[(k) * 2 + (m)]
    float c1 = 1.0f;
    CC(0, 0) = 1.0f;
    for (int n = 1; n < N; ++n) {
        const int mn = n < 1 ? n : 1;
        float c2 = 1.0f;
        for (int nu = 0; nu < n; ++nu) {
            const float c3 = x[n] - x[nu];
            c2 *= c3;
            if (nu == n - 1) {
                for (int m = mn; m >= 1; --m)
                    CC(n, m) = c1 / c2 * ((float)m * CC(n - 1, m - 1) - (x[n] - xi) * CC(n - 1, m));
                CC(n, 0) = c1 / c2 * (-(x[n] - xi)) * CC(n - 1, 0);
            }
            for (int m = mn; m >= 1; --m)
                CC(nu, m) = ((x[n] - xi) * CC(nu, m) - (float)m * CC(nu, m - 1)) / c3;
            CC(nu, 0) = (x[n] - xi) * CC(nu, 0) / c3;
        }
        c1 = c2;
    }
    for (int k = 0; k < N; ++k) w[k] = CC(k, 1);
#undef CC
}

static float amo_proj(const Amo* a, int ci, const float* d)
{
    return (float)(a->centers[3 * ci] * d[0] + a->centers[3 * ci + 1] * d[1]
                   + a->centers[3 * ci + 2] * d[2]);
}

/* Dominant axis of a direction if its largest |component| > 0.9, else -1
 * (acoustic_amr.cpp:351-357, :465-471). */
static int amo_axis(float ax, float ay, float az)
{
    float best = ax > ay ? ax : ay;
    best = best > az ? best : az;
    if (!(best > 0.9f)) return -1;
    if (ax >= ay && ax >= az) return 0;
    if (ay >= ax && ay >= az) return 1;
    return 2;
}

/* amr_build_dir_neighbors, acoustic_amr.cpp:329-391: per axis and side the
 * unique axis-aligned neighbour, -1 if none or conflicting. */
static void amo_build_dir_neighbors(Amo* a)
{
    uint8_t* st[6];
    for (int q = 0; q < 6; ++q) {
        a->dnb[q] = (int*)malloc(sizeof(int) * (size_t)a->nc);
        st[q] = (uint8_t*)calloc((size_t)a->nc, 1);
        for (int i = 0; i < a->nc; ++i) a->dnb[q][i] = -1;
    }
    for (int f = 0; f < a->nf; ++f) {
        const int cn = a->neg[f], cp = a->pos[f];
        const float dx = (float)(a->centers[3 * cp] - a->centers[3 * cn]);
        const float dy = (float)(a->centers[3 * cp + 1] - a->centers[3 * cn + 1]);
        const float dz = (float)(a->centers[3 * cp + 2] - a->centers[3 * cn + 2]);
        const float len = sqrtf(dx * dx + dy * dy + dz * dz);
        if (len < 1e-15f) continue;
        const int axis = amo_axis(fabsf(dx / len), fabsf(dy / len), fabsf(dz / len));
        if (axis < 0) continue;
        const float sd = axis == 0 ? dx : (axis == 1 ? dy : dz);
        /* table 2*axis holds the + neighbour, 2*axis+1 the - neighbour */
        const int lo = sd > 0.0f ? cn : cp, hi = sd > 0.0f ? cp : cn;
        int* tp = &a->dnb[2 * axis][lo];
        uint8_t* sp = &st[2 * axis][lo];
        if (*sp == 0) { *tp = hi; *sp = 1; } else if (*sp == 1 && *tp != hi) { *tp = -1; *sp = 2; }
        int* tn = &a->dnb[2 * axis + 1][hi];
        uint8_t* sn = &st[2 * axis + 1][hi];
        if (*sn == 0) { *tn = lo; *sn = 1; } else if (*sn == 1 && *tn != lo) { *tn = -1; *sn = 2; }
    }
    for (int q = 0; q < 6; ++q) free(st[q]);
}

/* amr_walk_stencil, acoustic_amr.cpp:444-547 */
static int amo_walk(const Amo* a, int seed, float fproj, const float* cd, const float* wd,
                    int depth, int* oc, float* ox)
{
    oc[0] = seed;
    ox[0] = amo_proj(a, seed, cd) - fproj;
    int found = 1, cur = seed;
    const int axis = amo_axis(fabsf(wd[0]), fabsf(wd[1]), fabsf(wd[2]));
    if (axis >= 0) {
        const int* tbl = a->dnb[2 * axis + (wd[axis] > 0.0f ? 0 : 1)];
        while (found < depth) {
            const int nb = tbl[cur];
            if (nb < 0 || nb >= a->nc) break;
            if (a->type[nb] == 1 || a->type[nb] == 2) break;
            oc[found] = nb;
            ox[found] = amo_proj(a, nb, cd) - fproj;
            ++found;
            cur = nb;
        }
        return found;
    }
    while (found < depth) {
        int best_next = -1;
        float best_dot = 0.5f;
        for (int k = a->csr_start[cur]; k < a->csr_start[cur + 1]; ++k) {
            const int fk = a->csr_face[k];
            const float sg = (a->csr_w[k] > 0.f) ? 1.f : -1.f;
            const int fn = a->neg[fk], fp = a->pos[fk];
            const float dx = (float)(a->centers[3 * fp] - a->centers[3 * fn]);
            const float dy = (float)(a->centers[3 * fp + 1] - a->centers[3 * fn + 1]);
            const float dz = (float)(a->centers[3 * fp + 2] - a->centers[3 * fn + 2]);
            const float len = sqrtf(dx * dx + dy * dy + dz * dz);
            if (len < 1e-15f) continue;
            const float dot = sg * (dx * wd[0] + dy * wd[1] + dz * wd[2]) / len;
            if (dot > best_dot) {
                const int nb = (sg > 0.f) ? fp : fn;
                if (nb >= 0 && nb < a->nc && a->type[nb] != 1 && a->type[nb] != 2) {
                    best_dot = dot;
                    best_next = nb;
                }
            }
        }
        if (best_next < 0) break;
        oc[found] = best_next;
        ox[found] = amo_proj(a, best_next, cd) - fproj;
        ++found;
        cur = best_next;
    }
    return found;
}

/* amr_build_stencil_range, acoustic_amr.cpp:549-629: 4 cells on each side,
 * Fornberg weights when both sides are complete, else the two-point
 * difference; unused slots hold cell 0 with weight 0. */
static void amo_build_stencil(Amo* a)
{
    amo_build_dir_neighbors(a);
    a->sc = (int*)calloc((size_t)a->nf * 8, sizeof(int));
    a->sw = (float*)calloc((size_t)a->nf * 8, sizeof(float));
    #pragma omp parallel for schedule(dynamic, 4096)
    for (int f = 0; f < a->nf; ++f) {
        const int cn = a->neg[f], cp = a->pos[f];
        const float dxf = (float)(a->centers[3 * cp] - a->centers[3 * cn]);
        const float dyf = (float)(a->centers[3 * cp + 1] - a->centers[3 * cn + 1]);
        const float dzf = (float)(a->centers[3 * cp + 2] - a->centers[3 * cn + 2]);
        const float len = sqrtf(dxf * dxf + dyf * dyf + dzf * dzf);
        if (!(len >= 1e-15f)) continue;
        const float dir[3] = {dxf / len, dyf / len, dzf / len};
        const float mdir[3] = {-dir[0], -dir[1], -dir[2]};
        const float fproj = 0.5f * (amo_proj(a, cn, dir) + amo_proj(a, cp, dir));
        int nc_[4], pc_[4];
        float nx_[4], px_[4];
        const int nn = amo_walk(a, cn, fproj, dir, mdir, 4, nc_, nx_);
        const int np = amo_walk(a, cp, fproj, dir, dir, 4, pc_, px_);
        if (nn == 4 && np == 4) {
            float xs[8], w[8];
            int cs[8];
            for (int k = 0; k < 4; ++k) { cs[k] = nc_[3 - k]; xs[k] = nx_[3 - k]; }
            for (int k = 0; k < 4; ++k) { cs[4 + k] = pc_[k]; xs[4 + k] = px_[k]; }
            amo_fornberg_d1(xs, 8, 0.0f, w);
            for (int k = 0; k < 4; ++k) {
                a->sc[8 * (size_t)f + (3 - k)] = cs[k];
                a->sw[8 * (size_t)f + (3 - k)] = w[k];
                a->sc[8 * (size_t)f + 4 + k] = cs[4 + k];
                a->sw[8 * (size_t)f + 4 + k] = w[4 + k];
            }
        } else if (nn + np >= 2) {
            const float inv_d = a->inv_d[f];  /* face_distance > 0 is validated */
            a->sc[8 * (size_t)f] = cn;
            a->sw[8 * (size_t)f] = -inv_d;
            a->sc[8 * (size_t)f + 4] = cp;
            a->sw[8 * (size_t)f + 4] = inv_d;
        }
    }
}

SK_EXPORT int amo_get_stencil(const Amo* a, int* sc, float* sw)
{
    if (!a || !a->sc) return AMO_ERR_NULL;
    memcpy(sc, a->sc, sizeof(int) * 8 * (size_t)a->nf);
    memcpy(sw, a->sw, sizeof(float) * 8 * (size_t)a->nf);
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
    free(a->volumes); free(a->face_area); free(a->face_open);
    free(a->sc); free(a->sw);
    for (int q = 0; q < 6; ++q) free(a->dnb[q]);
    free(a->w); free(a->w_prev); free(a->w_new); free(a->ext); free(a->active);
    free(a->ab_start); free(a->ab_face); free(a->ab_side); free(a->ab_wgt);
    free(a->bl_start); free(a->bl_face); free(a->bl_side); free(a->bl_wgt);
    free(a->cell_above); free(a->cell_below);
    free(a->bridge_idx); free(a->neck_idx); free(a->bridge_wgt); free(a->neck_wgt);
    for (int m = 0; m < a->n_mics; ++m) {
        free(a->mp_cell[m]); free(a->mp_w[m]);
        free(a->mv_face[m]); free(a->mv_wx[m]); free(a->mv_wy[m]); free(a->mv_wz[m]);
    }
    free(a->mp_n); free(a->mp_cell); free(a->mp_w);
    free(a->mv_n); free(a->mv_face); free(a->mv_wx); free(a->mv_wy); free(a->mv_wz);
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
    if (order != 2 && order != 8) return NULL;
    Amo* a = (Amo*)calloc(1, sizeof(Amo));
    a->nc = nc; a->nf = nf; a->order = order;
    a->c = c; a->rho = rho;
    a->dt = ((order == 8) ? 0.10 : 0.77) * min_dx / (c * sqrt(3.0));  /* :744-745 */
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
    a->volumes = (double*)malloc(sizeof(double) * (size_t)nc);
    memcpy(a->volumes, vol, sizeof(double) * (size_t)nc);
    a->face_area = (double*)malloc(sizeof(double) * (size_t)nf);
    memcpy(a->face_area, fa, sizeof(double) * (size_t)nf);
    a->face_open = (double*)malloc(sizeof(double) * (size_t)nf);
    memcpy(a->face_open, fo, sizeof(double) * (size_t)nf);
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
    if (order == 8) amo_build_stencil(a);                             /* :862-869 */
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
    if (a->plate_on) {
        memset(a->w, 0, sizeof(float) * (size_t)a->n_pnodes);
        memset(a->w_prev, 0, sizeof(float) * (size_t)a->n_pnodes);
        memset(a->ext, 0, sizeof(float) * (size_t)a->n_pnodes);
    }
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

/* plate_biharmonic, acoustic_amr.cpp:939-954: 13-point stencil, zero
 * outside the plate grid. */
static float amo_biharm(const float* w, int i, int j, int Nx, int Ny, float idx4)
{
#define W_(ii, jj) (((ii) < 0 || (ii) >= Nx || (jj) < 0 || (jj) >= Ny) ? 0.0f : w[(ii) * Ny + (jj)])
    float L = W_(i - 2, j) + W_(i + 2, j) + W_(i, j - 2) + W_(i, j + 2)
            + 2.0f * (W_(i - 1, j - 1) + W_(i - 1, j + 1) + W_(i + 1, j - 1) + W_(i + 1, j + 1))
            - 8.0f * (W_(i - 1, j) + W_(i + 1, j) + W_(i, j - 1) + W_(i, j + 1))
            + 20.0f * W_(i, j);
#undef W_
    return L * idx4;
}

/* amr_apply_plate_bc, acoustic_amr.cpp:1050-1072 */
static void amo_plate_bc(Amo* a, float dt)
{
    for (int n = 0; n < a->n_active; ++n) {
        const int flat = a->active[n];
        const float vp = (a->w[flat] - a->w_prev[flat]) / dt;
        for (int k = a->ab_start[n]; k < a->ab_start[n + 1]; ++k) a->v[a->ab_face[k]] = vp;
        for (int k = a->bl_start[n]; k < a->bl_start[n + 1]; ++k) a->v[a->bl_face[k]] = -vp;
    }
}

/* amr_plate_step, acoustic_amr.cpp:956-1047 */
static void amo_plate_step(Amo* a, float dt)
{
    const int Nx = a->pNx, Ny = a->pNy;
    const float dx = a->p_dx;
    const float dx4 = 1.0f / (dx * dx * dx * dx);
    const float dt2 = dt * dt, inv_rh = 1.0f / a->p_rh;
    const float damp_fwd = 1.0f + 0.5f * a->p_aM * dt;
    const float damp_bwd = 1.0f - 0.5f * a->p_aM * dt;
    const float cD0 = a->p_D * (1.0f + a->p_bK / dt);
    const float cDp = a->p_D * (a->p_bK / dt);
    memset(a->w_new, 0, sizeof(float) * (size_t)a->n_pnodes);
    for (int n = 0; n < a->n_active; ++n) {
        const int flat = a->active[n], i = flat / Ny, j = flat % Ny;
        float Pa = 0.0f, Pb = 0.0f;
        if (a->ab_start[n + 1] > a->ab_start[n]) {
            for (int k = a->ab_start[n]; k < a->ab_start[n + 1]; ++k)
                Pa += a->ab_wgt[k] * a->p[a->ab_side[k]];
        } else if (a->cell_above[n] >= 0) {
            Pa = a->p[a->cell_above[n]];
        }
        if (a->bl_start[n + 1] > a->bl_start[n]) {
            for (int k = a->bl_start[n]; k < a->bl_start[n + 1]; ++k)
                Pb += a->bl_wgt[k] * a->p[a->bl_side[k]];
        } else if (a->cell_below[n] >= 0) {
            Pb = a->p[a->cell_below[n]];
        }
        const float F_acou = (Pb - Pa) * dx * dx;
        const float F_ext = a->ext[flat];
        const float L4w = amo_biharm(a->w, i, j, Nx, Ny, dx4);
        const float L4wp = amo_biharm(a->w_prev, i, j, Nx, Ny, dx4);
        const float rhs = F_acou + F_ext - cD0 * L4w + cDp * L4wp;
        a->w_new[flat] = (damp_bwd * (2.0f * a->w[flat] - a->w_prev[flat]) + dt2 * inv_rh * rhs)
                         / damp_fwd;
    }
    for (int n = 0; n < a->n_active; ++n) {
        const int flat = a->active[n];
        a->w_prev[flat] = a->w[flat];
        a->w[flat] = a->w_new[flat];
    }
    memset(a->ext, 0, sizeof(float) * (size_t)a->n_pnodes);
}

/* amr_setup_plate, acoustic_amr.cpp:1412-1531 */
SK_EXPORT int amo_setup_plate(Amo* a, int Nx, int Ny, float pdx, const float* origin,
                              const uint8_t* act, float rh, float D, float aM, float bK,
                              int n_active, const int32_t* flat_idx,
                              const int32_t* ab_start, const int32_t* ab_idx, const float* ab_w,
                              const int32_t* bl_start, const int32_t* bl_idx, const float* bl_w,
                              const int32_t* c_above, const int32_t* c_below)
{
    (void)origin;
    if (!a) return AMO_ERR_NULL;
    if (!act || !flat_idx || !ab_start || !bl_start || !c_above || !c_below) return AMO_ERR_NULL;
    if (Nx <= 0 || Ny <= 0 || pdx <= 0.0f) return AMO_ERR_DIM;
    if (rh <= 0.0f || D <= 0.0f) return AMO_ERR_DIM;
    if (n_active < 0) return AMO_ERR_DIM;
    const int nn = Nx * Ny;
    int cnt = 0;
    for (int k = 0; k < nn; ++k) if (act[k]) ++cnt;
    if (cnt != n_active) return AMO_ERR_DIM;
    const int na = ab_start[n_active], nb = bl_start[n_active];
    for (int k = 0; k < na; ++k) if (ab_idx[k] < 0 || ab_idx[k] >= a->nf) return AMO_ERR_DIM;
    for (int k = 0; k < nb; ++k) if (bl_idx[k] < 0 || bl_idx[k] >= a->nf) return AMO_ERR_DIM;
    for (int n = 0; n < n_active; ++n) {
        if (c_above[n] != -1 && (c_above[n] < 0 || c_above[n] >= a->nc)) return AMO_ERR_DIM;
        if (c_below[n] != -1 && (c_below[n] < 0 || c_below[n] >= a->nc)) return AMO_ERR_DIM;
    }
    a->pNx = Nx; a->pNy = Ny; a->p_dx = pdx; a->n_pnodes = nn;
    a->p_rh = rh; a->p_D = D; a->p_aM = aM; a->p_bK = bK;
    free(a->w); free(a->w_prev); free(a->w_new); free(a->ext); free(a->active);
    free(a->ab_start); free(a->ab_face); free(a->ab_side); free(a->ab_wgt);
    free(a->bl_start); free(a->bl_face); free(a->bl_side); free(a->bl_wgt);
    free(a->cell_above); free(a->cell_below);
    a->w = (float*)calloc((size_t)nn, sizeof(float));
    a->w_prev = (float*)calloc((size_t)nn, sizeof(float));
    a->w_new = (float*)calloc((size_t)nn, sizeof(float));
    a->ext = (float*)calloc((size_t)nn, sizeof(float));
    a->n_active = n_active;
    a->active = (int*)malloc(sizeof(int) * ((size_t)n_active + 1));
    memcpy(a->active, flat_idx, sizeof(int) * (size_t)n_active);
    a->ab_start = (int*)malloc(sizeof(int) * ((size_t)n_active + 1));
    a->bl_start = (int*)malloc(sizeof(int) * ((size_t)n_active + 1));
    for (int n = 0; n <= n_active; ++n) {
        a->ab_start[n] = ab_start[n] - ab_start[0];
        a->bl_start[n] = bl_start[n] - bl_start[0];
    }
    const int la = ab_start[n_active] - ab_start[0], lb = bl_start[n_active] - bl_start[0];
    a->ab_face = (int*)malloc(sizeof(int) * ((size_t)la + 1));
    a->ab_side = (int*)malloc(sizeof(int) * ((size_t)la + 1));
    a->ab_wgt = (float*)malloc(sizeof(float) * ((size_t)la + 1));
    a->bl_face = (int*)malloc(sizeof(int) * ((size_t)lb + 1));
    a->bl_side = (int*)malloc(sizeof(int) * ((size_t)lb + 1));
    a->bl_wgt = (float*)malloc(sizeof(float) * ((size_t)lb + 1));
    for (int k = 0; k < la; ++k) {
        a->ab_face[k] = ab_idx[ab_start[0] + k];
        a->ab_wgt[k] = ab_w[ab_start[0] + k];
        a->ab_side[k] = a->pos[a->ab_face[k]];      /* air cell above: pos side */
    }
    for (int k = 0; k < lb; ++k) {
        a->bl_face[k] = bl_idx[bl_start[0] + k];
        a->bl_wgt[k] = bl_w[bl_start[0] + k];
        a->bl_side[k] = a->neg[a->bl_face[k]];      /* air cell below: neg side */
    }
    a->cell_above = (int*)malloc(sizeof(int) * ((size_t)n_active + 1));
    a->cell_below = (int*)malloc(sizeof(int) * ((size_t)n_active + 1));
    memcpy(a->cell_above, c_above, sizeof(int) * (size_t)n_active);
    memcpy(a->cell_below, c_below, sizeof(int) * (size_t)n_active);
    for (int n = 0; n < n_active; ++n) {                  /* every node coupled */
        if (a->ab_start[n + 1] == a->ab_start[n] && a->cell_above[n] < 0) return AMO_ERR_DIM;
        if (a->bl_start[n + 1] == a->bl_start[n] && a->cell_below[n] < 0) return AMO_ERR_DIM;
    }
    a->plate_on = 1;
    return AMO_OK;
}

/* amr_set_bridge_plate_sources / amr_set_neck_plate_sources, :1534-1563 */
SK_EXPORT int amo_set_bridge_sources(Amo* a, int n, const int32_t* idx, const float* wgt)
{
    if (!a) return AMO_ERR_NULL;
    if (!a->plate_on) return AMO_ERR_DIM;
    if (n <= 0 || !idx || !wgt) return AMO_ERR_DIM;
    for (int i = 0; i < n; ++i) if (idx[i] < 0 || idx[i] >= a->n_pnodes) return AMO_ERR_DIM;
    free(a->bridge_idx); free(a->bridge_wgt);
    a->bridge_idx = (int*)malloc(sizeof(int) * (size_t)n);
    a->bridge_wgt = (float*)malloc(sizeof(float) * (size_t)n);
    memcpy(a->bridge_idx, idx, sizeof(int) * (size_t)n);
    memcpy(a->bridge_wgt, wgt, sizeof(float) * (size_t)n);
    a->n_bridge = n;
    return AMO_OK;
}

SK_EXPORT int amo_set_neck_sources(Amo* a, int n, const int32_t* idx, const float* wgt)
{
    if (!a) return AMO_ERR_NULL;
    if (n == 0) { a->n_neck = 0; return AMO_OK; }
    if (!a->plate_on) return AMO_ERR_DIM;
    if (!idx || !wgt) return AMO_ERR_DIM;
    for (int i = 0; i < n; ++i) if (idx[i] < 0 || idx[i] >= a->n_pnodes) return AMO_ERR_DIM;
    free(a->neck_idx); free(a->neck_wgt);
    a->neck_idx = (int*)malloc(sizeof(int) * (size_t)n);
    a->neck_wgt = (float*)malloc(sizeof(float) * (size_t)n);
    memcpy(a->neck_idx, idx, sizeof(int) * (size_t)n);
    memcpy(a->neck_wgt, wgt, sizeof(float) * (size_t)n);
    a->n_neck = n;
    return AMO_OK;
}

/* amr_inject_bridge_drive / amr_inject_plate_force, :1567-1597 */
SK_EXPORT int amo_inject_bridge_drive(Amo* a, const float* drives, int n, float scale)
{
    if (!a || !drives) return AMO_ERR_NULL;
    if (!a->plate_on) return AMO_ERR_DIM;
    if (n != a->n_bridge) return AMO_ERR_DIM;
    for (int i = 0; i < n; ++i)
        a->ext[a->bridge_idx[i]] += drives[i] * a->bridge_wgt[i] * scale;
    return AMO_OK;
}

SK_EXPORT int amo_inject_plate_force(Amo* a, int n, const int32_t* idx, const float* wgt,
                                     float total)
{
    if (!a) return AMO_ERR_NULL;
    if (!a->plate_on) return AMO_ERR_DIM;
    if (n <= 0 || !idx || !wgt) return AMO_ERR_DIM;
    for (int i = 0; i < n; ++i) {
        if (idx[i] < 0 || idx[i] >= a->n_pnodes) return AMO_ERR_DIM;
        a->ext[idx[i]] += wgt[i] * total;
    }
    return AMO_OK;
}

/* plate samplers and getter, :1600-1644 */
SK_EXPORT int amo_sample_plate_batch(const Amo* a, int n, const int* idx4, const float* wgt4,
                                     float* out)
{
    if (!a || !idx4 || !wgt4 || !out) return AMO_ERR_NULL;
    if (!a->plate_on) return AMO_ERR_DIM;
    for (int q = 0; q < n; ++q) {
        float v = 0.0f;
        for (int k = 0; k < 4; ++k) {
            const int f = idx4[q * 4 + k];
            if (f < 0 || f >= a->n_pnodes) return AMO_ERR_DIM;
            v += wgt4[q * 4 + k] * a->w[f];
        }
        out[q] = v;
    }
    return AMO_OK;
}

SK_EXPORT int amo_sample_plate_weighted(const Amo* a, int n, const int* idx, const float* wgt,
                                        float* out)
{
    if (!a || !idx || !wgt || !out) return AMO_ERR_NULL;
    if (!a->plate_on) return AMO_ERR_DIM;
    float v = 0.0f;
    for (int i = 0; i < n; ++i) {
        if (idx[i] < 0 || idx[i] >= a->n_pnodes) return AMO_ERR_DIM;
        v += wgt[i] * a->w[idx[i]];
    }
    *out = v;
    return AMO_OK;
}

SK_EXPORT int amo_get_plate(const Amo* a, float* out, int n)
{
    if (!a || !out) return AMO_ERR_NULL;
    if (!a->plate_on) return AMO_ERR_DIM;
    if (n != a->n_pnodes) return AMO_ERR_DIM;
    memcpy(out, a->w, sizeof(float) * (size_t)n);
    return AMO_OK;
}

static void amo_push(int** ia, float** fa, int n, int iv, float fv)
{
    *ia = (int*)realloc(*ia, sizeof(int) * ((size_t)n + 1));
    *fa = (float*)realloc(*fa, sizeof(float) * ((size_t)n + 1));
    (*ia)[n] = iv;
    (*fa)[n] = fv;
}

/* amr_setup_mic_samplers, acoustic_amr.cpp:1647-1760 */
SK_EXPORT int amo_setup_mics(Amo* a, int n_mics, const float* pos, const float* axis,
                             int need_v)
{
    if (!a || !pos || !axis) return AMO_ERR_NULL;
    if (n_mics <= 0) return AMO_ERR_DIM;
    if (n_mics > a->n_mics) {                              /* resize: keep, add empty */
        a->mp_n = (int*)realloc(a->mp_n, sizeof(int) * (size_t)n_mics);
        a->mp_cell = (int**)realloc(a->mp_cell, sizeof(int*) * (size_t)n_mics);
        a->mp_w = (float**)realloc(a->mp_w, sizeof(float*) * (size_t)n_mics);
        a->mv_n = (int*)realloc(a->mv_n, sizeof(int) * (size_t)n_mics);
        a->mv_face = (int**)realloc(a->mv_face, sizeof(int*) * (size_t)n_mics);
        a->mv_wx = (float**)realloc(a->mv_wx, sizeof(float*) * (size_t)n_mics);
        a->mv_wy = (float**)realloc(a->mv_wy, sizeof(float*) * (size_t)n_mics);
        a->mv_wz = (float**)realloc(a->mv_wz, sizeof(float*) * (size_t)n_mics);
        for (int m = a->n_mics; m < n_mics; ++m) {
            a->mp_n[m] = 0; a->mp_cell[m] = NULL; a->mp_w[m] = NULL;
            a->mv_n[m] = 0; a->mv_face[m] = NULL;
            a->mv_wx[m] = NULL; a->mv_wy[m] = NULL; a->mv_wz[m] = NULL;
        }
    } else {
        for (int m = n_mics; m < a->n_mics; ++m) {        /* resize: drop the tail */
            free(a->mp_cell[m]); free(a->mp_w[m]);
            free(a->mv_face[m]); free(a->mv_wx[m]); free(a->mv_wy[m]); free(a->mv_wz[m]);
        }
    }
    a->n_mics = n_mics;
    double mean_vol = 0.0;
    float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int t = 0; t < a->n_acoustic; ++t) {
        const int idx = a->acoustic[t];
        mean_vol += a->volumes[idx];
        for (int d = 0; d < 3; ++d) {
            const float c = (float)a->centers[3 * idx + d];
            if (c < lo[d]) lo[d] = c;
            if (c > hi[d]) hi[d] = c;
        }
    }
    if (a->n_acoustic > 0) mean_vol /= (double)a->n_acoustic;
    const float R = 2.0f * (float)cbrt(mean_vol);
    for (int m = 0; m < n_mics; ++m) {
        float q[3];
        for (int d = 0; d < 3; ++d) {
            const float x = pos[3 * m + d];
            q[d] = fmaxf(lo[d], fminf(hi[d], x));
        }
        float wsum = 0.0f;
        for (int t = 0; t < a->n_acoustic; ++t) {
            const int idx = a->acoustic[t];
            const float dx = q[0] - (float)a->centers[3 * idx + 0];
            const float dy = q[1] - (float)a->centers[3 * idx + 1];
            const float dz = q[2] - (float)a->centers[3 * idx + 2];
            const float r = sqrtf(dx * dx + dy * dy + dz * dz);
            if (r < R) {
                const float w = (r < 1e-9f) ? 1e9f : 1.0f / r;
                amo_push(&a->mp_cell[m], &a->mp_w[m], a->mp_n[m]++, idx, w);
                wsum += w;
            }
        }
        if (a->mp_n[m] == 0) return AMO_ERR_DIM;
        const float inv = 1.0f / wsum;
        for (int k = 0; k < a->mp_n[m]; ++k) a->mp_w[m][k] *= inv;
        if (need_v) {
            float tot = 0.0f;
            const int k0 = a->mv_n[m];
            for (int f = 0; f < a->nf; ++f) {
                const int p0 = a->neg[f], p1 = a->pos[f];
                const float fx = 0.5f * (float)(a->centers[3 * p0] + a->centers[3 * p1]);
                const float fy = 0.5f * (float)(a->centers[3 * p0 + 1] + a->centers[3 * p1 + 1]);
                const float fz = 0.5f * (float)(a->centers[3 * p0 + 2] + a->centers[3 * p1 + 2]);
                const float ex = fx - q[0], ey = fy - q[1], ez = fz - q[2];
                const float r2 = sqrtf(ex * ex + ey * ey + ez * ez);
                if (r2 >= R || a->face_open[f] < 0.5) continue;
                float nx = (float)(a->centers[3 * p1] - a->centers[3 * p0]);
                float ny = (float)(a->centers[3 * p1 + 1] - a->centers[3 * p0 + 1]);
                float nz = (float)(a->centers[3 * p1 + 2] - a->centers[3 * p0 + 2]);
                const float nl = sqrtf(nx * nx + ny * ny + nz * nz);
                if (nl < 1e-12f) continue;
                nx /= nl; ny /= nl; nz /= nl;
                const float aw = (float)(a->face_area[f] * a->face_open[f]);
                const int k = a->mv_n[m]++;
                a->mv_face[m] = (int*)realloc(a->mv_face[m], sizeof(int) * ((size_t)k + 1));
                a->mv_wx[m] = (float*)realloc(a->mv_wx[m], sizeof(float) * ((size_t)k + 1));
                a->mv_wy[m] = (float*)realloc(a->mv_wy[m], sizeof(float) * ((size_t)k + 1));
                a->mv_wz[m] = (float*)realloc(a->mv_wz[m], sizeof(float) * ((size_t)k + 1));
                a->mv_face[m][k] = f;
                a->mv_wx[m][k] = nx * aw;
                a->mv_wy[m][k] = ny * aw;
                a->mv_wz[m][k] = nz * aw;
                tot += aw;
            }
            (void)k0;
            if (tot > 1e-20f) {
                const float inv_a = 1.0f / tot;
                for (int k = 0; k < a->mv_n[m]; ++k) {
                    a->mv_wx[m][k] *= inv_a; a->mv_wy[m][k] *= inv_a; a->mv_wz[m][k] *= inv_a;
                }
            }
        }
    }
    return AMO_OK;
}

/* amr_sample_mics, acoustic_amr.cpp:1762-1792 */
SK_EXPORT int amo_sample_mics(const Amo* a, int n_mics, float* op, float* ovx, float* ovy,
                              float* ovz)
{
    if (!a || !op) return AMO_ERR_NULL;
    if (n_mics != a->n_mics) return AMO_ERR_DIM;
    for (int m = 0; m < n_mics; ++m) {
        float p = 0.0f;
        for (int k = 0; k < a->mp_n[m]; ++k) p += a->mp_w[m][k] * a->p[a->mp_cell[m][k]];
        op[m] = p;
        if (ovx && ovy && ovz) {
            float vx = 0.0f, vy = 0.0f, vz = 0.0f;
            for (int k = 0; k < a->mv_n[m]; ++k) {
                const float vf = a->v[a->mv_face[m][k]];
                vx += a->mv_wx[m][k] * vf;
                vy += a->mv_wy[m][k] * vf;
                vz += a->mv_wz[m][k] * vf;
            }
            ovx[m] = vx; ovy[m] = vy; ovz[m] = vz;
        }
    }
    return AMO_OK;
}

/* amr_step, order-2 branch, acoustic_amr.cpp:1100-1204. */
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
        if (a->order == 8) {
            #pragma omp parallel for schedule(static)
            for (int f = 0; f < nf; ++f) {                            /* :1127-1143 */
                float g = 0.0f;
                for (int k = 0; k < 8; ++k) g += a->sw[8 * (size_t)f + k] * a->p[a->sc[8 * (size_t)f + k]];
                a->v[f] = (a->v[f] - dt_over_rho * g) * a->vdamp[f];
            }
        } else {
            #pragma omp parallel for schedule(static)
            for (int f = 0; f < nf; ++f) {                            /* :1147-1154 */
                const float g = (a->p[a->pos[f]] - a->p[a->neg[f]]) * a->inv_d[f];
                a->v[f] = (a->v[f] - dt_over_rho * g) * a->vdamp[f];
            }
        }
        if (a->plate_on) amo_plate_bc(a, dt);                         /* :1157 */
        #pragma omp parallel for schedule(static)
        for (int c = 0; c < nc; ++c) {                                /* :1163-1172 */
            float d = 0.0f;
            for (int k = a->csr_start[c]; k < a->csr_start[c + 1]; ++k)
                d += a->csr_w[k] * a->v[a->csr_face[k]];
            a->p[c] = a->p[c] * a->pdamp[c] - bulk * d * a->inv_den[c] * a->psrc[c];
        }
        for (int t = 0; t < a->n_wallplate; ++t) a->p[a->wallplate[t]] = 0.0f;  /* :1174-1176 */
        if (a->plate_on) amo_plate_step(a, dt);                        /* :1179 */
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
