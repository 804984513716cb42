/*
 * fdtd_oracle_impl.h — body of the CPU oracle, compiled once per precision.
 *
 * TEST INFRASTRUCTURE ONLY.  This restates, in plain C, the reference's
 * uniform-grid FDTD step (csrc/kernels/acoustic_fdtd.cpp) so that tests and
 * the CPU baseline can check the HIP engine against it.  Nothing in the
 * product library links or calls this code.
 *
 * The including file defines:
 *   REAL   float or double (arithmetic and field storage type)
 *   SUF(x) name suffixing macro (x##_f32 / x##_f64)
 *
 * Float evaluation order mirrors the reference expression by expression so
 * the float32 build reproduces the reference's rounding; the file must be
 * compiled with -ffp-contract=off (no FMA contraction), as the reference is
 * built without FMA (x86-64 baseline, no -march).
 */

typedef struct SUF(OrcState) {
    int Nx, Ny, Nz, N, Nplate;
    int plate_iz;           /* validated value (acoustic_fdtd.cpp:359)       */
    int order, boundary, n_pml, hetero;
    float dt_f;             /* dt as the reference computes it, in float     */
    REAL dx, dt, c, rho_air, inv_dx, vel_coeff, bulk_coeff;
    REAL C1, C2;
    REAL *P, *Ps;           /* P^n and the scratch buffer (holds P^{n-1})    */
    uint8_t* ct;            /* cell types after PML retyping and plate mask  */
    REAL *alpha, *pdamp;    /* per cell                                     */
    REAL *vx, *vy, *vz;     /* packed staggered layouts                     */
    REAL *vxd, *vyd, *vzd;  /* per-face damping multipliers                 */
    REAL *fx, *fy, *fz;     /* per-face open fractions                      */
    int has_frac;
    int *cfx, *cfy, *cfz;   /* closed-face lists (frac == 0), :297-323        */
    int ncfx, ncfy, ncfz;
    REAL *rho, *K, *kap;    /* per-cell density, bulk, K*dt/dx (hetero)      */
    REAL *bu;               /* per-cell buoyancy 1/rho (hetero)              */
    REAL dtdx;              /* dt/dx                                         */
    REAL gside[6];          /* (1-R)/(1+R) per domain side                   */
    /* Kirchhoff plate */
    REAL *w, *wp, *wt, *ext;
    uint8_t* pa;
    int nact;
    int *act, *above, *below, *b13, *vzb, *vza;
    REAL plate_a, plate_bh, plate_vd, plate_bk;
    /* bridge sources */
    int nbsrc; int* bidx; REAL* bwgt;
    /* point sources */
    int npsrc, plen; int* pidx; float* psig;
    /* receivers */
    int nrec; int* ridx; REAL* rwgt;
    double* trace; int trace_cap, trace_n;
    int step_count, phase;
} SUF(OrcState);

#define OST SUF(OrcState)

static int SUF(idx3)(const OST* s, int i, int j, int k) { return k + s->Nz * (j + s->Ny * i); }
static int SUF(solid)(uint8_t c) { return c == 1 || c == 2; }

/* PML profile — acoustic_fdtd.cpp:216-222 */
static REAL SUF(pml_sigma)(int d, int n_pml, REAL dt)
{
    if (n_pml <= 0 || d <= 0) return (REAL)0;
    REAL frac = (REAL)d / (REAL)n_pml;
    REAL sigma_max = (REAL)3.45f / (dt * (REAL)n_pml);
    return sigma_max * frac * frac * frac;
}

SK_EXPORT void SUF(orc_destroy)(OST* s)
{
    if (!s) return;
    free(s->P); free(s->Ps); free(s->ct); free(s->alpha); free(s->pdamp);
    free(s->vx); free(s->vy); free(s->vz); free(s->vxd); free(s->vyd); free(s->vzd);
    free(s->fx); free(s->fy); free(s->fz); free(s->rho); free(s->K); free(s->kap); free(s->bu);
    free(s->w); free(s->wp); free(s->wt); free(s->ext); free(s->pa);
    free(s->act); free(s->above); free(s->below); free(s->b13); free(s->vzb); free(s->vza);
    free(s->bidx); free(s->bwgt); free(s->pidx); free(s->psig);
    free(s->ridx); free(s->rwgt); free(s->trace);
    free(s->cfx); free(s->cfy); free(s->cfz);
    free(s);
}

/* Closed-face worklists (faces with frac == 0), as build_closed_face_lists
 * (acoustic_fdtd.cpp:297-323); rebuilt whenever fractions change. */
static int SUF(build_closed)(OST* s)
{
    int n[3] = {(s->Nx - 1) * s->Ny * s->Nz, s->Nx * (s->Ny - 1) * s->Nz,
                s->Nx * s->Ny * (s->Nz - 1)};
    REAL* F[3] = {s->fx, s->fy, s->fz};
    int** L[3] = {&s->cfx, &s->cfy, &s->cfz};
    int* C[3] = {&s->ncfx, &s->ncfy, &s->ncfz};
    for (int a = 0; a < 3; ++a) {
        free(*L[a]); *L[a] = NULL; *C[a] = 0;
        int m = 0;
        for (int q = 0; q < n[a]; ++q) if (F[a][q] == (REAL)0) ++m;
        if (m == 0) continue;
        *L[a] = (int*)malloc((size_t)m * sizeof(int));
        if (!*L[a]) return -1;
        m = 0;
        for (int q = 0; q < n[a]; ++q) if (F[a][q] == (REAL)0) (*L[a])[m++] = q;
        *C[a] = m;
    }
    return 0;
}

/* Construction — acoustic_fdtd.cpp:327-768 plus the extension options. */
SK_EXPORT OST* SUF(orc_create)(const OrcConfig* cf, const uint8_t* cell_type,
                               const uint8_t* plate_active,
                               const float* rho_in, const float* bulk_in)
{
    int Nx = cf->Nx, Ny = cf->Ny, Nz = cf->Nz;
    if (Nx <= 0 || Ny <= 0 || Nz <= 0 || cf->dx <= 0.0f || cf->c <= 0.0f
        || !cell_type || !plate_active) return NULL;
    if (cf->order != 2 && cf->order != 4) return NULL;
    if (cf->boundary == 1 && cf->order != 2) return NULL;
    OST* s = (OST*)calloc(1, sizeof(OST));
    if (!s) return NULL;
    int N = Nx * Ny * Nz;
    s->Nx = Nx; s->Ny = Ny; s->Nz = Nz; s->N = N; s->Nplate = Nx * Ny;
    s->order = cf->order; s->boundary = cf->boundary; s->n_pml = cf->n_pml;
    s->hetero = (rho_in != NULL) || (bulk_in != NULL);
    s->plate_iz = (cf->plate_iz >= 0 && cf->plate_iz < Nz) ? cf->plate_iz : Nz / 2;

    /* dt: the reference rule in float (acoustic_fdtd.cpp:348), or a Courant
     * number over the fastest cell (extension). */
    float dt_f;
    if (cf->courant <= 0.0f) {
        dt_f = 0.77f * cf->dx / (cf->c * 1.732050808f);
    } else {
        double cmax = cf->c;
        if (s->hetero) {
            cmax = 0.0;
            for (int q = 0; q < N; ++q) {
                double r = rho_in ? rho_in[q] : cf->rho_air;
                double kk = bulk_in ? bulk_in[q] : (double)((cf->rho_air * cf->c) * cf->c);
                double cc = sqrt(kk / r);
                if (cc > cmax) cmax = cc;
            }
        }
        dt_f = (float)((double)cf->courant * (double)cf->dx / cmax);
    }
    s->dt_f = dt_f;
    s->dt = (REAL)dt_f;
    s->dx = (REAL)cf->dx; s->c = (REAL)cf->c; s->rho_air = (REAL)cf->rho_air;
    s->inv_dx = (REAL)1 / s->dx;
    s->dtdx = s->dt / s->dx;
    s->vel_coeff = s->dt / (s->rho_air * s->dx);                       /* :1089 */
    s->bulk_coeff = s->rho_air * s->c * s->c * s->dt * s->inv_dx;        /* :1176 */
    s->C1 = (REAL)9 / (REAL)8;                                          /* :889 */
    s->C2 = (REAL)-1 / (REAL)24;                                        /* :890 */

    s->P = (REAL*)calloc(N, sizeof(REAL));
    s->Ps = (REAL*)calloc(N, sizeof(REAL));
    s->ct = (uint8_t*)malloc(N);
    s->alpha = (REAL*)calloc(N, sizeof(REAL));
    s->pdamp = (REAL*)calloc(N, sizeof(REAL));
    if (!s->P || !s->Ps || !s->ct || !s->alpha || !s->pdamp) goto fail;
    memcpy(s->ct, cell_type, N);

    /* PML depth and retyping — acoustic_fdtd.cpp:377-396 */
    {
        int np = cf->n_pml;
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz; ++k) {
                    int di = (i < np) ? (np - i) : ((i >= Nx - np) ? (i - (Nx - np) + 1) : 0);
                    int dj = (j < np) ? (np - j) : ((j >= Ny - np) ? (j - (Ny - np) + 1) : 0);
                    int dk = (k < np) ? (np - k) : ((k >= Nz - np) ? (k - (Nz - np) + 1) : 0);
                    int d = di > dj ? di : dj;
                    if (dk > d) d = dk;
                    REAL a = SUF(pml_sigma)(d, np, s->dt);
                    int q = SUF(idx3)(s, i, j, k);
                    s->alpha[q] = a;
                    if (s->ct[q] == 0 && a > (REAL)0) s->ct[q] = 3;
                }
    }

    /* Plate mask on the pressure grid — acoustic_fdtd.cpp:413-424 */
    s->pa = (uint8_t*)malloc(s->Nplate);
    s->w = (REAL*)calloc(s->Nplate, sizeof(REAL));
    s->wp = (REAL*)calloc(s->Nplate, sizeof(REAL));
    s->wt = (REAL*)calloc(s->Nplate, sizeof(REAL));
    s->ext = (REAL*)calloc(s->Nplate, sizeof(REAL));
    if (!s->pa || !s->w || !s->wp || !s->wt || !s->ext) goto fail;
    memcpy(s->pa, plate_active, s->Nplate);
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j) {
            int g = SUF(idx3)(s, i, j, s->plate_iz);
            if (s->pa[j + Ny * i]) {
                if (s->ct[g] == 0 || s->ct[g] == 3) s->ct[g] = 2;
            } else if (s->ct[g] == 2) {
                s->ct[g] = 0;
            }
        }

    /* Plate constants — acoustic_fdtd.cpp:435-447 */
    {
        REAL dt = s->dt, dx = s->dx;
        REAL m = (REAL)cf->plate_mass, D = (REAL)cf->plate_D;
        s->plate_a = (dt * dt) / m;
        REAL dx4 = dx * dx * dx * dx;
        s->plate_bh = D * (dt * dt) / (m * dx4);
        s->plate_vd = (REAL)1 - (REAL)cf->plate_alpha_M * dt;
        s->plate_bk = (dt > (REAL)0) ? ((REAL)cf->plate_beta_K / dt * s->plate_bh) : (REAL)0;
    }

    /* Velocities, damping, fractions — acoustic_fdtd.cpp:449-561 */
    {
        int nvx = (Nx - 1) * Ny * Nz, nvy = Nx * (Ny - 1) * Nz, nvz = Nx * Ny * (Nz - 1);
        int mx = nvx > 0 ? nvx : 1, my = nvy > 0 ? nvy : 1, mz = nvz > 0 ? nvz : 1;
        s->vx = (REAL*)calloc(mx, sizeof(REAL)); s->vy = (REAL*)calloc(my, sizeof(REAL));
        s->vz = (REAL*)calloc(mz, sizeof(REAL));
        s->vxd = (REAL*)calloc(mx, sizeof(REAL)); s->vyd = (REAL*)calloc(my, sizeof(REAL));
        s->vzd = (REAL*)calloc(mz, sizeof(REAL));
        s->fx = (REAL*)calloc(mx, sizeof(REAL)); s->fy = (REAL*)calloc(my, sizeof(REAL));
        s->fz = (REAL*)calloc(mz, sizeof(REAL));
        if (!s->vx || !s->vy || !s->vz || !s->vxd || !s->vyd || !s->vzd
            || !s->fx || !s->fy || !s->fz) goto fail;
        const uint8_t* ct = s->ct;
        REAL dt = s->dt;
        for (int i = 0; i < Nx - 1; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz; ++k) {
                    int v = k + Nz * (j + Ny * i);
                    int L = SUF(idx3)(s, i, j, k), R = SUF(idx3)(s, i + 1, j, k);
                    s->fx[v] = (SUF(solid)(ct[L]) || SUF(solid)(ct[R])) ? (REAL)0 : (REAL)1;
                    s->vxd[v] = (REAL)1 - (REAL)0.5f * (s->alpha[L] + s->alpha[R]) * dt;
                }
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny - 1; ++j)
                for (int k = 0; k < Nz; ++k) {
                    int v = k + Nz * (j + (Ny - 1) * i);
                    int L = SUF(idx3)(s, i, j, k), R = SUF(idx3)(s, i, j + 1, k);
                    s->fy[v] = (SUF(solid)(ct[L]) || SUF(solid)(ct[R])) ? (REAL)0 : (REAL)1;
                    s->vyd[v] = (REAL)1 - (REAL)0.5f * (s->alpha[L] + s->alpha[R]) * dt;
                }
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz - 1; ++k) {
                    int v = k + (Nz - 1) * (j + Ny * i);
                    int L = SUF(idx3)(s, i, j, k), R = SUF(idx3)(s, i, j, k + 1);
                    if (k == s->plate_iz - 1 || k == s->plate_iz)
                        s->fz[v] = (REAL)1;     /* plate-layer faces stay open (:527-528) */
                    else
                        s->fz[v] = (SUF(solid)(ct[L]) || SUF(solid)(ct[R])) ? (REAL)0 : (REAL)1;
                    s->vzd[v] = (REAL)1 - (REAL)0.5f * (s->alpha[L] + s->alpha[R]) * dt;
                }
    }
    for (int q = 0; q < N; ++q) {                                        /* :469-476 */
        REAL den = (REAL)1 + s->alpha[q] * s->dt;
        s->pdamp[q] = (REAL)1 / den;
    }

    /* Heterogeneous media (extension): per-cell rho, K and K*dt/dx. */
    if (s->hetero) {
        s->rho = (REAL*)malloc(N * sizeof(REAL));
        s->K = (REAL*)malloc(N * sizeof(REAL));
        s->kap = (REAL*)malloc(N * sizeof(REAL));
        s->bu = (REAL*)malloc(N * sizeof(REAL));
        if (!s->rho || !s->K || !s->kap || !s->bu) goto fail;
        REAL K0 = s->rho_air * s->c * s->c;
        for (int q = 0; q < N; ++q) {
            s->rho[q] = rho_in ? (REAL)rho_in[q] : s->rho_air;
            s->K[q] = bulk_in ? (REAL)bulk_in[q] : K0;
            s->kap[q] = s->K[q] * s->dt * s->inv_dx;
            s->bu[q] = (REAL)1 / s->rho[q];
        }
    }
    for (int f = 0; f < 6; ++f) {
        REAL R = (REAL)cf->refl[f];
        s->gside[f] = ((REAL)1 - R) / ((REAL)1 + R);
    }

    /* Plate precompute — acoustic_fdtd.cpp:671-752 */
    {
        static const int DI[13] = {0, 1, -1, 0, 0, 1, 1, -1, -1, 2, -2, 0, 0};
        static const int DJ[13] = {0, 0, 0, 1, -1, 1, -1, 1, -1, 0, 0, 2, -2};
        int na = 0;
        for (int p = 0; p < s->Nplate; ++p) if (plate_active[p]) ++na;
        s->nact = na;
        int m = na > 0 ? na : 1;
        s->act = (int*)malloc(m * sizeof(int)); s->above = (int*)malloc(m * sizeof(int));
        s->below = (int*)malloc(m * sizeof(int)); s->b13 = (int*)malloc(13 * m * sizeof(int));
        s->vzb = (int*)malloc(m * sizeof(int)); s->vza = (int*)malloc(m * sizeof(int));
        if (!s->act || !s->above || !s->below || !s->b13 || !s->vzb || !s->vza) goto fail;
        int n = 0, piz = s->plate_iz;
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) {
                int p = j + Ny * i;
                if (!plate_active[p]) continue;
                s->act[n] = p;
                s->above[n] = (piz + 1 < Nz) ? SUF(idx3)(s, i, j, piz + 1) : -1;
                s->below[n] = (piz - 1 >= 0) ? SUF(idx3)(s, i, j, piz - 1) : -1;
                for (int t = 0; t < 13; ++t) {
                    int ii = i + DI[t], jj = j + DJ[t];
                    if (ii < 0 || ii >= Nx || jj < 0 || jj >= Ny || !plate_active[jj + Ny * ii]) {
                        ii = i; jj = j;
                    }
                    s->b13[n * 13 + t] = jj + Ny * ii;
                }
                if (piz > 0) {
                    int b = SUF(idx3)(s, i, j, piz - 1);
                    s->vzb[n] = SUF(solid)(s->ct[b]) ? -1 : (piz - 1) + (Nz - 1) * (j + Ny * i);
                } else s->vzb[n] = -1;
                if (piz < Nz - 1) {
                    int a = SUF(idx3)(s, i, j, piz + 1);
                    s->vza[n] = SUF(solid)(s->ct[a]) ? -1 : piz + (Nz - 1) * (j + Ny * i);
                } else s->vza[n] = -1;
                ++n;
            }
    }
    if (SUF(build_closed)(s)) goto fail;
    return s;
fail:
    SUF(orc_destroy)(s);
    return NULL;
}

SK_EXPORT float SUF(orc_get_dt)(const OST* s) { return s->dt_f; }
SK_EXPORT int SUF(orc_get_step_count)(const OST* s) { return s->step_count; }

/* acoustic_fdtd.cpp:1064-1077 (fractions in packed layouts; NULL = unchanged) */
SK_EXPORT int SUF(orc_set_face_fractions)(OST* s, const float* fx, const float* fy,
                                          const float* fz)
{
    int nvx = (s->Nx - 1) * s->Ny * s->Nz, nvy = s->Nx * (s->Ny - 1) * s->Nz;
    int nvz = s->Nx * s->Ny * (s->Nz - 1);
    if (fx) for (int q = 0; q < nvx; ++q) s->fx[q] = (REAL)fx[q];
    if (fy) for (int q = 0; q < nvy; ++q) s->fy[q] = (REAL)fy[q];
    if (fz) for (int q = 0; q < nvz; ++q) s->fz[q] = (REAL)fz[q];
    if (fx || fy || fz) s->has_frac = 1;
    return SUF(build_closed)(s);
}

/* acoustic_fdtd.cpp:798-830 */
SK_EXPORT int SUF(orc_set_bridge_sources)(OST* s, int n, const int* idx, const float* wgt)
{
    if (n <= 0 || !idx || !wgt) return -1;
    float wsum = 0.0f;
    for (int i = 0; i < n; ++i) wsum += fabsf(wgt[i]);
    float wscale = (wsum > 1e-12f) ? 1.0f / wsum : 1.0f;
    free(s->bidx); free(s->bwgt);
    s->bidx = (int*)malloc(n * sizeof(int));
    s->bwgt = (REAL*)malloc(n * sizeof(REAL));
    for (int i = 0; i < n; ++i) { s->bidx[i] = idx[i]; s->bwgt[i] = (REAL)(wgt[i] * wscale); }
    s->nbsrc = n;
    return 0;
}

/* acoustic_fdtd.cpp:1699-1721 */
SK_EXPORT int SUF(orc_inject_bridge_drive)(OST* s, const float* drives, float force_scale)
{
    if (!s->bidx || s->nbsrc == 0) return 0;
    if (!drives) return -1;
    REAL inv_dx2 = (REAL)1 / (s->dx * s->dx);
    for (int i = 0; i < s->nbsrc; ++i) {
        int p = s->bidx[i] / s->Nz;
        s->ext[p] += (REAL)force_scale * (REAL)drives[i] * inv_dx2;
    }
    return 0;
}

/* acoustic_fdtd.cpp:1723-1742 */
SK_EXPORT int SUF(orc_inject_plate_force)(OST* s, int n, const int* pidx, const float* wgt,
                                          float total)
{
    if (n <= 0 || total == 0.0f) return 0;
    REAL inv_dx2 = (REAL)1 / (s->dx * s->dx);
    for (int i = 0; i < n; ++i) {
        int p = pidx[i];
        if (p < 0 || p >= s->Nplate || !s->pa[p]) continue;
        s->ext[p] += (REAL)total * (REAL)wgt[i] * inv_dx2;
    }
    return 0;
}

/* Point pressure sources (extension; additive as acoustic_amr.cpp:1227-1244). */
SK_EXPORT int SUF(orc_set_point_sources)(OST* s, int n, const int* idx, const float* sig, int len)
{
    free(s->pidx); free(s->psig); s->pidx = NULL; s->psig = NULL; s->npsrc = 0; s->plen = 0;
    if (n <= 0) return 0;
    for (int i = 0; i < n; ++i)
        if (idx[i] < 0 || idx[i] >= s->N || SUF(solid)(s->ct[idx[i]])) return -2;
    s->pidx = (int*)malloc(n * sizeof(int));
    s->psig = (float*)malloc((size_t)n * len * sizeof(float));
    memcpy(s->pidx, idx, n * sizeof(int));
    memcpy(s->psig, sig, (size_t)n * len * sizeof(float));
    s->npsrc = n; s->plen = len;
    return 0;
}

/* Trilinear 8-tap stencil — acoustic_fdtd.cpp:1567-1596 */
static void SUF(p_sampler)(const OST* s, const float* xyz, int* idx, REAL* wgt)
{
    int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    REAL gx = (REAL)xyz[0], gy = (REAL)xyz[1], gz = (REAL)xyz[2];
    if (gx < (REAL)0) gx = (REAL)0; if (gx > (REAL)(Nx - 1)) gx = (REAL)(Nx - 1);
    if (gy < (REAL)0) gy = (REAL)0; if (gy > (REAL)(Ny - 1)) gy = (REAL)(Ny - 1);
    if (gz < (REAL)0) gz = (REAL)0; if (gz > (REAL)(Nz - 1)) gz = (REAL)(Nz - 1);
    int i0 = (int)gx, i1 = (i0 < Nx - 1) ? i0 + 1 : i0;
    int j0 = (int)gy, j1 = (j0 < Ny - 1) ? j0 + 1 : j0;
    int k0 = (int)gz, k1 = (k0 < Nz - 1) ? k0 + 1 : k0;
    REAL fx = gx - (REAL)i0, fy = gy - (REAL)j0, fz = gz - (REAL)k0;
    REAL one = (REAL)1;
    idx[0] = SUF(idx3)(s, i0, j0, k0); wgt[0] = (one - fx) * (one - fy) * (one - fz);
    idx[1] = SUF(idx3)(s, i1, j0, k0); wgt[1] = (fx) * (one - fy) * (one - fz);
    idx[2] = SUF(idx3)(s, i0, j1, k0); wgt[2] = (one - fx) * (fy) * (one - fz);
    idx[3] = SUF(idx3)(s, i1, j1, k0); wgt[3] = (fx) * (fy) * (one - fz);
    idx[4] = SUF(idx3)(s, i0, j0, k1); wgt[4] = (one - fx) * (one - fy) * (fz);
    idx[5] = SUF(idx3)(s, i1, j0, k1); wgt[5] = (fx) * (one - fy) * (fz);
    idx[6] = SUF(idx3)(s, i0, j1, k1); wgt[6] = (one - fx) * (fy) * (fz);
    idx[7] = SUF(idx3)(s, i1, j1, k1); wgt[7] = (fx) * (fy) * (fz);
}

SK_EXPORT int SUF(orc_precompute_pressure_samplers)(const OST* s, int n, const float* xyz,
                                                    int* idx8, float* wgt8)
{
    for (int r = 0; r < n; ++r) {
        REAL w[8];
        SUF(p_sampler)(s, xyz + 3 * r, idx8 + 8 * r, w);
        for (int t = 0; t < 8; ++t) wgt8[8 * r + t] = (float)w[t];
    }
    return 0;
}

/* Sampling in the precomputed form's summation order — acoustic_fdtd.cpp:1672-1673 */
static REAL SUF(gather8)(const REAL* F, const int* id, const REAL* w)
{
    return w[0] * F[id[0]] + w[1] * F[id[1]] + w[2] * F[id[2]] + w[3] * F[id[3]]
         + w[4] * F[id[4]] + w[5] * F[id[5]] + w[6] * F[id[6]] + w[7] * F[id[7]];
}

SK_EXPORT int SUF(orc_set_receivers)(OST* s, int n, const float* xyz)
{
    free(s->ridx); free(s->rwgt); s->ridx = NULL; s->rwgt = NULL; s->nrec = 0;
    if (n <= 0) return 0;
    s->ridx = (int*)malloc(8 * n * sizeof(int));
    s->rwgt = (REAL*)malloc(8 * n * sizeof(REAL));
    for (int r = 0; r < n; ++r) SUF(p_sampler)(s, xyz + 3 * r, s->ridx + 8 * r, s->rwgt + 8 * r);
    s->nrec = n;
    return 0;
}

SK_EXPORT int SUF(orc_record_receivers)(OST* s, int max_steps)
{
    free(s->trace);
    s->trace = (double*)calloc((size_t)max_steps * (s->nrec > 0 ? s->nrec : 1), sizeof(double));
    s->trace_cap = max_steps; s->trace_n = 0;
    return s->trace ? 0 : -1;
}

SK_EXPORT int SUF(orc_get_traces)(const OST* s, double* out, int len)
{
    int n = s->trace_n * s->nrec;
    if (len < n) return -2;
    memcpy(out, s->trace, (size_t)n * sizeof(double));
    return s->trace_n;
}

/* Kirchhoff plate — acoustic_fdtd.cpp:984-1053 */
static int SUF(plate_step)(OST* s)
{
    static const REAL W13[13] = {20, -8, -8, -8, -8, 2, 2, 2, 2, 1, 1, 1, 1};
    const REAL* w = s->w; const REAL* wp = s->wp; REAL* wt = s->wt;
    for (int n = 0; n < s->nact; ++n) {
        int p = s->act[n];
        const int* nb = s->b13 + 13 * n;
        REAL bh = (REAL)0;
        for (int t = 0; t < 13; ++t) bh += W13[t] * w[nb[t]];
        REAL Pa = (s->above[n] >= 0) ? s->P[s->above[n]] : (REAL)0;
        REAL Pb = (s->below[n] >= 0) ? s->P[s->below[n]] : (REAL)0;
        REAL fb = s->ext[p];
        REAL bhp = (REAL)0;
        for (int t = 0; t < 13; ++t) bhp += W13[t] * wp[nb[t]];
        REAL vp = w[p] - wp[p];
        REAL wn = w[p] + s->plate_vd * vp + s->plate_a * ((Pb - Pa) + fb)
                - s->plate_bh * bh - s->plate_bk * (bh - bhp);
        if (!isfinite((double)wn) || fabs((double)wn) > (double)0.02f) return -3;
        wt[p] = wn;
    }
    memset(s->ext, 0, s->Nplate * sizeof(REAL));
    REAL* t = s->wp; s->wp = s->w; s->w = s->wt; s->wt = t;
    return 0;
}

/* Divergence of the staggered velocity at cell (i,j,k): the reference's
 * zero-padded 4th-order form (acoustic_fdtd.cpp:894-952) or the 2nd-order
 * Yee form.  fr selects the face-fraction–weighted variant. */
static REAL SUF(divergence)(const OST* s, int i, int j, int k, int fr)
{
    int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    const REAL *X = s->vx, *Y = s->vy, *Z = s->vz;
    const REAL *FX = s->fx, *FY = s->fy, *FZ = s->fz;
#define AX(ii) (fr ? X[k + Nz * (j + Ny * (ii))] * FX[k + Nz * (j + Ny * (ii))] : X[k + Nz * (j + Ny * (ii))])
#define AY(jj) (fr ? Y[k + Nz * ((jj) + (Ny - 1) * i)] * FY[k + Nz * ((jj) + (Ny - 1) * i)] : Y[k + Nz * ((jj) + (Ny - 1) * i)])
#define AZ(kk) (fr ? Z[(kk) + (Nz - 1) * (j + Ny * i)] * FZ[(kk) + (Nz - 1) * (j + Ny * i)] : Z[(kk) + (Nz - 1) * (j + Ny * i)])
    REAL zero = (REAL)0;
    REAL xp = (i < Nx - 1) ? AX(i) : zero, xm = (i > 0) ? AX(i - 1) : zero;
    REAL yp = (j < Ny - 1) ? AY(j) : zero, ym = (j > 0) ? AY(j - 1) : zero;
    REAL zp = (k < Nz - 1) ? AZ(k) : zero, zm = (k > 0) ? AZ(k - 1) : zero;
    if (s->order == 2)
        return (xp - xm) + (yp - ym) + (zp - zm);
    REAL xpp = (i < Nx - 2) ? AX(i + 1) : zero, xmm = (i > 1) ? AX(i - 2) : zero;
    REAL ypp = (j < Ny - 2) ? AY(j + 1) : zero, ymm = (j > 1) ? AY(j - 2) : zero;
    REAL zpp = (k < Nz - 2) ? AZ(k + 1) : zero, zmm = (k > 1) ? AZ(k - 2) : zero;
#undef AX
#undef AY
#undef AZ
    return s->C1 * ((xp - xm) + (yp - ym) + (zp - zm))
         + s->C2 * ((xpp - xmm) + (ypp - ymm) + (zpp - zmm));
}

/* Face gradient: 4th order where both outer cells exist, else 2nd order
 * (acoustic_fdtd.cpp:1106-1112); always 2nd order for the Yee stencil. */
static REAL SUF(gradient)(const OST* s, int Lm, int L, int R, int Rp, int inner)
{
    const REAL* P = s->P;
    if (s->order == 4 && inner)
        return s->C1 * (P[R] - P[L]) + s->C2 * (P[Rp] - P[Lm]);
    return P[R] - P[L];
}

/* Velocity coefficient of the face between cells L and R: dt/(rho dx) for a
 * homogeneous medium (acoustic_fdtd.cpp:1089); for a heterogeneous medium
 * (extension) the face buoyancy is the mean of the cell buoyancies 1/rho,
 * i.e. the face density is the harmonic mean of the cell densities. */
static REAL SUF(face_coeff)(const OST* s, int L, int R)
{
    if (!s->hetero) return s->vel_coeff;
    return ((REAL)0.5f * (s->bu[L] + s->bu[R])) * s->dtdx;
}

/* One leapfrog update — acoustic_fdtd.cpp:1081-1241 (+ extension terms). */
static void SUF(pressure_step)(OST* s)
{
    int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz, sx = Ny * Nz, sy = Nz;
    /* Velocity update (:1099-1145); x slabs in parallel as FDTDPool (:1147) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx - 1; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                int v = k + Nz * (j + Ny * i);
                int L = SUF(idx3)(s, i, j, k), R = L + sx;
                int inner = (i > 0 && i < Nx - 2);
                REAL g = SUF(gradient)(s, inner ? L - sx : L, L, R, inner ? R + sx : R, inner);
                s->vx[v] = (s->vx[v] - SUF(face_coeff)(s, L, R) * g) * s->vxd[v];
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny - 1; ++j)
            for (int k = 0; k < Nz; ++k) {
                int v = k + Nz * (j + (Ny - 1) * i);
                int L = SUF(idx3)(s, i, j, k), R = L + sy;
                int inner = (j > 0 && j < Ny - 2);
                REAL g = SUF(gradient)(s, inner ? L - sy : L, L, R, inner ? R + sy : R, inner);
                s->vy[v] = (s->vy[v] - SUF(face_coeff)(s, L, R) * g) * s->vyd[v];
            }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz - 1; ++k) {
                int v = k + (Nz - 1) * (j + Ny * i);
                int L = SUF(idx3)(s, i, j, k), R = L + 1;
                int inner = (k > 0 && k < Nz - 2);
                REAL g = SUF(gradient)(s, inner ? L - 1 : L, L, R, inner ? R + 1 : R, inner);
                s->vz[v] = (s->vz[v] - SUF(face_coeff)(s, L, R) * g) * s->vzd[v];
            }
    /* Closed faces (frac == 0) are zeroed from the worklists (:1152-1154) */
    for (int q = 0; q < s->ncfx; ++q) s->vx[s->cfx[q]] = (REAL)0;
    for (int q = 0; q < s->ncfy; ++q) s->vy[s->cfy[q]] = (REAL)0;
    for (int q = 0; q < s->ncfz; ++q) s->vz[s->cfz[q]] = (REAL)0;
    /* Plate velocity injection (:1158-1163) */
    for (int n = 0; n < s->nact; ++n) {
        int p = s->act[n];
        REAL vpl = (s->w[p] - s->wp[p]) * ((REAL)1 / s->dt);
        if (s->vzb[n] >= 0) s->vz[s->vzb[n]] = vpl;
        if (s->vza[n] >= 0) s->vz[s->vza[n]] = vpl;
    }
    /* Pressure update (:1176-1234): every non-solid cell; solid cells keep 0. */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j) {
            const int yee_run = s->order == 2 && !s->has_frac && i > 0 && i < Nx - 1
                              && j > 0 && j < Ny - 1;
            for (int k = 0; k < Nz; ++k) {
                int c = SUF(idx3)(s, i, j, k);
                if (SUF(solid)(s->ct[c])) continue;
                /* Interior cells of the Yee stencil without fractions: every
                 * face exists, so the divergence is the same three differences
                 * in the same order as divergence(), with direct offsets. */
                if (yee_run && k > 0 && k < Nz - 1) {
                    const REAL* X = s->vx; const REAL* Y = s->vy; const REAL* Z = s->vz;
                    int ox = c, oy = k + Nz * (j + (Ny - 1) * i), oz = k + (Nz - 1) * (j + Ny * i);
                    REAL dv = (X[ox] - X[ox - sx]) + (Y[oy] - Y[oy - Nz]) + (Z[oz] - Z[oz - 1]);
                    REAL kap = s->hetero ? s->kap[c] : s->bulk_coeff;
                    s->Ps[c] = (s->P[c] - kap * dv) * s->pdamp[c];
                    continue;
                }
                REAL dv = SUF(divergence)(s, i, j, k, s->has_frac);
                REAL kap = s->hetero ? s->kap[c] : s->bulk_coeff;
                int edge = (i == 0) | (i == Nx - 1) | (j == 0) | (j == Ny - 1)
                         | (k == 0) | (k == Nz - 1);
                if (s->boundary == 1 && edge) {
                    /* Locally reacting wall, semi-implicit in time (extension):
                     * v_b = Y * (p^{n+1}+p^n)/2 outward, Y = (1-R)/((1+R) rho c). */
                    REAL r = s->hetero ? s->rho[c] : s->rho_air;
                    REAL KK = s->hetero ? s->K[c] : s->rho_air * s->c * s->c;
                    REAL z = (REAL)sqrt((double)(r * KK));
                    REAL hk = (REAL)0.5f * kap;
                    REAL xi = (REAL)0;
                    if (i == 0) xi = xi + hk * (s->gside[0] / z);
                    if (i == Nx - 1) xi = xi + hk * (s->gside[1] / z);
                    if (j == 0) xi = xi + hk * (s->gside[2] / z);
                    if (j == Ny - 1) xi = xi + hk * (s->gside[3] / z);
                    if (k == 0) xi = xi + hk * (s->gside[4] / z);
                    if (k == Nz - 1) xi = xi + hk * (s->gside[5] / z);
                    REAL num = s->P[c] * ((REAL)1 - xi) - kap * dv;
                    s->Ps[c] = (num / ((REAL)1 + xi)) * s->pdamp[c];
                } else {
                    s->Ps[c] = (s->P[c] - kap * dv) * s->pdamp[c];
                }
            }
        }
    REAL* t = s->P; s->P = s->Ps; s->Ps = t;                            /* :1236-1239 */
}

/* acoustic_fdtd.cpp:1245-1268 */
static int SUF(check_stable)(OST* s)
{
    int start = s->phase;
    s->phase = (start + 1) % 1013;
    for (int q = start; q < s->N; q += 1013) {
        double p = (double)s->P[q];
        if (!isfinite(p) || p > 2e3 || p < -2e3) return -3;
    }
    return 0;
}

/* acoustic_fdtd.cpp:1272-1290, with point sources and receiver recording
 * (extension) after the pressure update. */
SK_EXPORT int SUF(orc_step)(OST* s, int n_steps)
{
    if (n_steps <= 0) return 0;
    for (int q = 0; q < n_steps; ++q) {
        int rc = SUF(plate_step)(s);
        if (rc) return rc;
        SUF(pressure_step)(s);
        for (int m = 0; m < s->npsrc; ++m)
            if (s->step_count < s->plen)
                s->P[s->pidx[m]] += (REAL)s->psig[(size_t)m * s->plen + s->step_count];
        if (s->trace && s->trace_n < s->trace_cap) {
            for (int r = 0; r < s->nrec; ++r)
                s->trace[(size_t)s->trace_n * s->nrec + r] =
                    (double)SUF(gather8)(s->P, s->ridx + 8 * r, s->rwgt + 8 * r);
            s->trace_n++;
        }
        ++s->step_count;
        if ((s->step_count & 0xFF) == 0)
            if (SUF(check_stable)(s)) return -3;
    }
    return 0;
}

/* acoustic_fdtd.cpp:1297-1310 */
SK_EXPORT int SUF(orc_reset)(OST* s)
{
    memset(s->P, 0, s->N * sizeof(REAL));
    memset(s->Ps, 0, s->N * sizeof(REAL));
    memset(s->w, 0, s->Nplate * sizeof(REAL)); memset(s->wp, 0, s->Nplate * sizeof(REAL));
    memset(s->ext, 0, s->Nplate * sizeof(REAL));
    memset(s->vx, 0, (s->Nx - 1) * s->Ny * s->Nz * sizeof(REAL));
    memset(s->vy, 0, s->Nx * (s->Ny - 1) * s->Nz * sizeof(REAL));
    memset(s->vz, 0, s->Nx * s->Ny * (s->Nz - 1) * sizeof(REAL));
    s->step_count = 0;
    s->trace_n = 0;
    return 0;
}

/* Field access in the reference layouts (acoustic_fdtd.cpp:1314-1430). */
SK_EXPORT int SUF(orc_get_fields)(const OST* s, REAL* P, REAL* vx, REAL* vy, REAL* vz)
{
    if (P) memcpy(P, s->P, s->N * sizeof(REAL));
    if (vx) memcpy(vx, s->vx, (s->Nx - 1) * s->Ny * s->Nz * sizeof(REAL));
    if (vy) memcpy(vy, s->vy, s->Nx * (s->Ny - 1) * s->Nz * sizeof(REAL));
    if (vz) memcpy(vz, s->vz, s->Nx * s->Ny * (s->Nz - 1) * sizeof(REAL));
    return 0;
}

/* Set P and packed velocities (test helper). */
SK_EXPORT int SUF(orc_set_fields)(OST* s, const REAL* P, const REAL* vx, const REAL* vy,
                                  const REAL* vz)
{
    if (P) memcpy(s->P, P, s->N * sizeof(REAL));
    if (vx) memcpy(s->vx, vx, (s->Nx - 1) * s->Ny * s->Nz * sizeof(REAL));
    if (vy) memcpy(s->vy, vy, s->Nx * (s->Ny - 1) * s->Nz * sizeof(REAL));
    if (vz) memcpy(s->vz, vz, s->Nx * s->Ny * (s->Nz - 1) * sizeof(REAL));
    return 0;
}

SK_EXPORT int SUF(orc_get_plate)(const OST* s, REAL* w)
{
    memcpy(w, s->w, s->Nplate * sizeof(REAL));
    return 0;
}

SK_EXPORT int SUF(orc_get_cell_type)(const OST* s, uint8_t* out)
{
    memcpy(out, s->ct, s->N);
    return 0;
}

/* acoustic_fdtd.cpp:1744-1760: weighted sum of displacements from 0, term by
 * term, skipping out-of-range indices. */
SK_EXPORT int SUF(orc_sample_plate_weighted)(const OST* s, int n, const int* pidx,
                                             const float* wgt, REAL* out)
{
    REAL acc = (REAL)0;
    for (int i = 0; i < n; ++i) {
        int p = pidx[i];
        if (p < 0 || p >= s->Nplate) continue;
        acc += s->w[p] * (REAL)wgt[i];
    }
    *out = acc;
    return 0;
}

/* acoustic_fdtd.cpp:1332-1353: four-tap gather per point, left to right. */
SK_EXPORT int SUF(orc_sample_plate_batch)(const OST* s, int n, const int* idx4,
                                          const float* wgt4, REAL* out)
{
    for (int q = 0; q < n; ++q) {
        const int* id = idx4 + 4 * q;
        const float* g = wgt4 + 4 * q;
        out[q] = (REAL)g[0] * s->w[id[0]] + (REAL)g[1] * s->w[id[1]]
               + (REAL)g[2] * s->w[id[2]] + (REAL)g[3] * s->w[id[3]];
    }
    return 0;
}

/* acoustic_fdtd.cpp:1355-1395 (summation starts from 0 as in the reference) */
SK_EXPORT int SUF(orc_sample_pressure)(const OST* s, int n, const float* xyz, REAL* out)
{
    for (int r = 0; r < n; ++r) {
        int id[8]; REAL w[8];
        SUF(p_sampler)(s, xyz + 3 * r, id, w);
        REAL p = (REAL)0;
        for (int t = 0; t < 8; ++t) p += w[t] * s->P[id[t]];
        out[r] = p;
    }
    return 0;
}

/* Staggered samplers — acoustic_fdtd.cpp:1598-1645 (axis with half-cell shift) */
static void SUF(v_sampler)(const OST* s, int ax, const float* xyz, int* idx, REAL* wgt)
{
    int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    int Nx1 = ax == 0 ? Nx - 1 : Nx, Ny1 = ax == 1 ? Ny - 1 : Ny, Nz1 = ax == 2 ? Nz - 1 : Nz;
    REAL half = (REAL)0.5f;
    REAL gx = (REAL)xyz[0] - (ax == 0 ? half : (REAL)0);
    REAL gy = (REAL)xyz[1] - (ax == 1 ? half : (REAL)0);
    REAL gz = (REAL)xyz[2] - (ax == 2 ? half : (REAL)0);
    if (gx < (REAL)0) gx = (REAL)0; if (gx > (REAL)(Nx1 - 1)) gx = (REAL)(Nx1 - 1);
    if (gy < (REAL)0) gy = (REAL)0; if (gy > (REAL)(Ny1 - 1)) gy = (REAL)(Ny1 - 1);
    if (gz < (REAL)0) gz = (REAL)0; if (gz > (REAL)(Nz1 - 1)) gz = (REAL)(Nz1 - 1);
    int i0 = (int)gx, i1 = (i0 < Nx1 - 1) ? i0 + 1 : i0;
    int j0 = (int)gy, j1 = (j0 < Ny1 - 1) ? j0 + 1 : j0;
    int k0 = (int)gz, k1 = (k0 < Nz1 - 1) ? k0 + 1 : k0;
    REAL fx = gx - (REAL)i0, fy = gy - (REAL)j0, fz = gz - (REAL)k0, one = (REAL)1;
#define VI(a, b, c) (ax == 2 ? ((c) + Nz1 * ((b) + Ny * (a))) : \
                     (ax == 1 ? ((c) + Nz * ((b) + Ny1 * (a))) : ((c) + Nz * ((b) + Ny * (a)))))
    idx[0] = VI(i0, j0, k0); wgt[0] = (one - fx) * (one - fy) * (one - fz);
    idx[1] = VI(i1, j0, k0); wgt[1] = (fx) * (one - fy) * (one - fz);
    idx[2] = VI(i0, j1, k0); wgt[2] = (one - fx) * (fy) * (one - fz);
    idx[3] = VI(i1, j1, k0); wgt[3] = (fx) * (fy) * (one - fz);
    idx[4] = VI(i0, j0, k1); wgt[4] = (one - fx) * (one - fy) * (fz);
    idx[5] = VI(i1, j0, k1); wgt[5] = (fx) * (one - fy) * (fz);
    idx[6] = VI(i0, j1, k1); wgt[6] = (one - fx) * (fy) * (fz);
    idx[7] = VI(i1, j1, k1); wgt[7] = (fx) * (fy) * (fz);
#undef VI
}

SK_EXPORT int SUF(orc_precompute_velocity_samplers)(const OST* s, int n, const float* xyz,
                                                    int* xi, float* xw, int* yi, float* yw,
                                                    int* zi, float* zw)
{
    int* I[3] = {xi, yi, zi}; float* W[3] = {xw, yw, zw};
    for (int ax = 0; ax < 3; ++ax)
        for (int r = 0; r < n; ++r) {
            REAL w[8];
            SUF(v_sampler)(s, ax, xyz + 3 * r, I[ax] + 8 * r, w);
            for (int t = 0; t < 8; ++t) W[ax][8 * r + t] = (float)w[t];
        }
    return 0;
}

/* acoustic_fdtd.cpp:1436-1538: summation from 0, term by term. */
SK_EXPORT int SUF(orc_sample_velocity)(const OST* s, int n, const float* xyz,
                                       REAL* ox, REAL* oy, REAL* oz)
{
    const REAL* F[3] = {s->vx, s->vy, s->vz};
    REAL* O[3] = {ox, oy, oz};
    for (int r = 0; r < n; ++r)
        for (int ax = 0; ax < 3; ++ax) {
            int id[8]; REAL w[8];
            SUF(v_sampler)(s, ax, xyz + 3 * r, id, w);
            REAL v = (REAL)0;
            for (int t = 0; t < 8; ++t) v += w[t] * F[ax][id[t]];
            O[ax][r] = v;
        }
    return 0;
}

/* FP64 energy terms: [0] sum P^n P^{n-1} / (2K) dV (uses the scratch buffer,
 * which holds P^{n-1} at non-solid cells), [1] sum rho_f v^2 / 2 dV over
 * open faces, [2] sum P^2/(2K) dV. */
SK_EXPORT int SUF(orc_energy)(const OST* s, double* out)
{
    double dV = (double)s->dx * (double)s->dx * (double)s->dx;
    double K0 = (double)s->rho_air * (double)s->c * (double)s->c;
    double e0 = 0, e1 = 0, e2 = 0;
    int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz;
    for (int q = 0; q < s->N; ++q) {
        if (SUF(solid)(s->ct[q])) continue;
        double K = s->hetero ? (double)s->K[q] : K0;
        e0 += (double)s->P[q] * (double)s->Ps[q] / (2.0 * K);
        e2 += (double)s->P[q] * (double)s->P[q] / (2.0 * K);
    }
    /* face density = 1 / face buoyancy (the weight that makes the discrete
     * energy of the heterogeneous Yee scheme exactly invariant) */
#define RHOF(L, R) (s->hetero ? 2.0 / ((double)s->bu[L] + (double)s->bu[R]) : (double)s->rho_air)
    for (int i = 0; i < Nx - 1; ++i) for (int j = 0; j < Ny; ++j) for (int k = 0; k < Nz; ++k) {
        int L = SUF(idx3)(s, i, j, k), R = SUF(idx3)(s, i + 1, j, k);
        double v = (double)s->vx[k + Nz * (j + Ny * i)];
        e1 += 0.5 * RHOF(L, R) * v * v;
    }
    for (int i = 0; i < Nx; ++i) for (int j = 0; j < Ny - 1; ++j) for (int k = 0; k < Nz; ++k) {
        int L = SUF(idx3)(s, i, j, k), R = SUF(idx3)(s, i, j + 1, k);
        double v = (double)s->vy[k + Nz * (j + (Ny - 1) * i)];
        e1 += 0.5 * RHOF(L, R) * v * v;
    }
    for (int i = 0; i < Nx; ++i) for (int j = 0; j < Ny; ++j) for (int k = 0; k < Nz - 1; ++k) {
        int L = SUF(idx3)(s, i, j, k), R = SUF(idx3)(s, i, j, k + 1);
        double v = (double)s->vz[k + (Nz - 1) * (j + Ny * i)];
        e1 += 0.5 * RHOF(L, R) * v * v;
    }
#undef RHOF
    out[0] = e0 * dV; out[1] = e1 * dV; out[2] = e2 * dV;
    return 0;
}

#undef OST
