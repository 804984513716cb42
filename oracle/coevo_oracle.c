/*
 * coevo_oracle.c — CPU oracle of the instrument co-evolver.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, as the checker.  The
 * product (libafdtd.so) never links, calls or falls back to it.
 *
 * A plain-C restatement of the reference's uniform co-evolver
 * (csrc/kernels/acoustic_coevolver.cpp): strings (two polarisations,
 * centred damped leapfrog, biharmonic stiffness, geometric tension), the
 * shared neck mode, per-string bridge impedances, bridge and neck loads on
 * the body plate, the body FDTD (the float32 FDTD oracle of fdtd_oracle.c,
 * compiled into this library), saddle feedback, modal projection, magnetic
 * and piezo pickups and polar microphones, per audio sample.  Each function
 * cites the reference lines it follows.
 *
 * Where the reference reduces through Eigen (geometric tension arc length
 * :563-593, modal GEMV :665-689, bridge-drive GEMV :2054-2062, pickup GEMV
 * :2152-2171) this oracle sums term by term in index order; Eigen is not
 * vendored in the reference (version unpinned), so the packet order of its
 * reductions is not a fixed target.  Element-wise Eigen expressions are
 * restated operation by operation in their expression-tree order (they
 * round identically).  Everything else follows the reference's own loops.
 *
 * Compiled with -ffp-contract=off (no FMA contraction), as the reference.
 */
#include "fdtd_oracle.c"
#include "amr_oracle.c"

#include <float.h>
#include <stdio.h>

#include "../include/afdtd_coevolver.h"

#define CO_N_MODES 32       /* acoustic_coevolver.cpp:40 */
#define CO_MAX_SUBSTEPS 16  /* :41 */
#define CO_OUTPUT_BUF 4096  /* :42 */
#define CO_MAX_SLOPE 0.25f  /* :43 */
#define CO_MAX_STRETCH 0.0015f
#define CO_PI 3.14159265358979323846

typedef struct {
    int onset, si;
    float pos, amp;
    int ndur;
} OPluck;

typedef struct {
    int n_segs, n_nodes;
    float *xyz, *tan, *nrm, *bin;
    float L, ds;
    float T0, T_eff, EA, mu, c_s, gamma, gamma_eff, dt_s, biharm;
    int n_sub;
    float *u[2], *up[2], *ut[2], *fe[2];
    float saddle_w, saddle_w_prev, bridge_w, bridge_w_prev;
    float b_m, b_k, b_c;
    float modal_re[CO_N_MODES], modal_im[CO_N_MODES];
    float* basis;
    float* fbuf; int nfbuf, fnode, fonset;
} OString;

typedef struct {
    int type; uint32_t mask; float sens;
    float** Bn; float** Bb;
} OPickup;

typedef struct {
    float g[3], axis[3], pa, pb, gain;
} OMic;

typedef struct OrcCoevo {
    int ns, npk, nmic;
    OString* s;
    OPickup* pk;
    OMic* mic;
    OrcState_f32* f;            /* uniform body, or */
    Amo* amr;                   /* AMR body */
    int Nx, Ny, Nz, plate_iz;
    float dx, origin[3], rho_air, c;
    int* bidx; float* bwgt; int nb;
    int* nidx; float* nwgt; int nn;
    float neck_w, neck_w_prev, neck_plate_w_prev;
    float n_m, n_k, n_c, nb_k, nb_c;
    float sr, dt_audio, dt_fdtd; int K;
    float gamma_scale;
    int modal_stride, modal_counter;
    float* pk_out; float* mic_out;   /* [npk][4096], [nmic][4096] */
    int out_pos, written;
    float* pk_acc;
    int* sad_idx; float* sad_wgt; float* sad_w;
    int* mP_i; float* mP_w;
    int* mV_i[3]; float* mV_w[3];
    int any_vel;
    float* sbk;                 /* [ns][nb] per-string bridge kernel */
    float* bdrive;
    int tot_seg; int* seg_off; int nmag; int* mag_map; float *K0, *K1, *v0f, *v1f, *mag_tmp;
    OPluck* ev; int nev, cap_ev, cursor;
    int cur_sample;
} OrcCoevo;

static void co_frame(OString* ss)                                   /* :325-396 */
{
    int N = ss->n_segs;
    ss->L = 0.0f;
    for (int s = 0; s < N; ++s) {
        float dx = ss->xyz[(s + 1) * 3 + 0] - ss->xyz[s * 3 + 0];
        float dy = ss->xyz[(s + 1) * 3 + 1] - ss->xyz[s * 3 + 1];
        float dz = ss->xyz[(s + 1) * 3 + 2] - ss->xyz[s * 3 + 2];
        float len = sqrtf(dx * dx + dy * dy + dz * dz);
        if (len < 1e-12f) len = 1e-12f;
        ss->L += len;
        ss->tan[s * 3 + 0] = dx / len;
        ss->tan[s * 3 + 1] = dy / len;
        ss->tan[s * 3 + 2] = dz / len;
    }
    ss->ds = ss->L / (float)N;
    float tx = ss->tan[0], ty = ss->tan[1], tz = ss->tan[2];
    float ref[3] = {0.0f, 0.0f, 0.0f};
    if (fabsf(ty) < 0.9f) ref[1] = 1.0f; else ref[2] = 1.0f;
    float dot = ref[0] * tx + ref[1] * ty + ref[2] * tz;
    float nx = ref[0] - dot * tx, ny = ref[1] - dot * ty, nz = ref[2] - dot * tz;
    float nl = sqrtf(nx * nx + ny * ny + nz * nz); if (nl < 1e-12f) nl = 1e-12f;
    nx /= nl; ny /= nl; nz /= nl;
    for (int s = 0; s < N; ++s) {
        float* t = ss->tan + s * 3;
        if (s > 0) {
            float* tp = ss->tan + (s - 1) * 3;
            float ax = tp[1] * t[2] - tp[2] * t[1];
            float ay = tp[2] * t[0] - tp[0] * t[2];
            float az = tp[0] * t[1] - tp[1] * t[0];
            float al = sqrtf(ax * ax + ay * ay + az * az);
            if (al > 1e-8f) {
                float th = asinf(fminf(al, 1.0f));
                float c_ = cosf(th), s_ = sinf(th) / al;
                ax *= s_; ay *= s_; az *= s_;
                float cx = ay * nz - az * ny, cy = az * nx - ax * nz, cz = ax * ny - ay * nx;
                float adn = ax * nx + ay * ny + az * nz;
                nx = c_ * nx + sinf(th) * cx + (1 - c_) * adn * ax;
                ny = c_ * ny + sinf(th) * cy + (1 - c_) * adn * ay;
                nz = c_ * nz + sinf(th) * cz + (1 - c_) * adn * az;
                dot = nx * t[0] + ny * t[1] + nz * t[2];
                nx -= dot * t[0]; ny -= dot * t[1]; nz -= dot * t[2];
                nl = sqrtf(nx * nx + ny * ny + nz * nz); if (nl < 1e-12f) nl = 1e-12f;
                nx /= nl; ny /= nl; nz /= nl;
            }
        }
        float* n = ss->nrm + s * 3;
        float* b = ss->bin + s * 3;
        n[0] = nx; n[1] = ny; n[2] = nz;
        b[0] = t[1] * n[2] - t[2] * n[1];
        b[1] = t[2] * n[0] - t[0] * n[2];
        b[2] = t[0] * n[1] - t[1] * n[0];
    }
}

static int co_string_init(OString* ss, const CoEvolverStringDef* d, float dt_fdtd)  /* :399-522 */
{
    int N = d->n_segs;
    if (N < 2 || !d->path_xyz || d->tension_N <= 0 || d->linear_mass_kgm <= 0) return -1;
    ss->n_segs = N; ss->n_nodes = N + 1;
    ss->xyz = (float*)calloc((N + 1) * 3, sizeof(float));
    ss->tan = (float*)calloc(N * 3, sizeof(float));
    ss->nrm = (float*)calloc(N * 3, sizeof(float));
    ss->bin = (float*)calloc(N * 3, sizeof(float));
    memcpy(ss->xyz, d->path_xyz, (N + 1) * 3 * sizeof(float));
    ss->T0 = d->tension_N; ss->T_eff = d->tension_N; ss->mu = d->linear_mass_kgm;
    ss->gamma = d->damping; ss->gamma_eff = d->damping;
    ss->EA = d->axial_stiffness_N;
    if (ss->EA <= 0.0f && ss->mu > 0.0f) ss->EA = (2.0e11f / 7850.0f) * ss->mu;
    ss->b_m = 0.00035f;
    {
        float om = 2.0f * (float)CO_PI * 450.0f;
        ss->b_k = ss->b_m * om * om;
        ss->b_c = ss->b_m * om / 2.5f;
    }
    co_frame(ss);
    ss->c_s = sqrtf(d->tension_N / d->linear_mass_kgm);
    {
        float EI = d->stiffness_EI, mu = d->linear_mass_kgm;
        float cst = (EI > 0.0f && mu > 0.0f)
            ? (EI / mu) * ((float)CO_PI / ss->ds) * ((float)CO_PI / ss->ds) : 0.0f;
        float c_eff = sqrtf(ss->c_s * ss->c_s + cst);
        float dtmax = 0.95f * ss->ds / c_eff;
        ss->n_sub = (int)ceilf(dt_fdtd / dtmax);
        if (ss->n_sub < 1) ss->n_sub = 1;
        ss->dt_s = dt_fdtd / (float)ss->n_sub;
        ss->biharm = (EI > 0.0f && mu > 0.0f)
            ? (EI / mu) * (ss->dt_s * ss->dt_s) / (ss->ds * ss->ds * ss->ds * ss->ds) : 0.0f;
    }
    for (int p = 0; p < 2; ++p) {
        ss->u[p] = (float*)calloc(N + 1, sizeof(float));
        ss->up[p] = (float*)calloc(N + 1, sizeof(float));
        ss->ut[p] = (float*)calloc(N + 1, sizeof(float));
        ss->fe[p] = (float*)calloc(N + 1, sizeof(float));
    }
    int Nn = ss->n_nodes;
    ss->basis = (float*)calloc((size_t)CO_N_MODES * Nn, sizeof(float));
    float inv = 1.0f / (float)(Nn > 1 ? Nn - 1 : 1);
    for (int k = 0; k < CO_N_MODES; ++k) {
        float kpi = (float)(k + 1) * (float)CO_PI;
        for (int i = 0; i < Nn; ++i) ss->basis[k * Nn + i] = sinf(kpi * (float)i * inv);
    }
    return 0;
}

static void co_string_free(OString* ss)
{
    for (int p = 0; p < 2; ++p) { free(ss->u[p]); free(ss->up[p]); free(ss->ut[p]); free(ss->fe[p]); }
    free(ss->xyz); free(ss->tan); free(ss->nrm); free(ss->bin); free(ss->basis); free(ss->fbuf);
}

/* Pickup B-kernels — :711-775 */
static float co_line_dist(const float* p, const float* a, const float* d)
{
    float ax = p[0] - a[0], ay = p[1] - a[1], az = p[2] - a[2];
    float pr = ax * d[0] + ay * d[1] + az * d[2];
    float rx = ax - pr * d[0], ry = ay - pr * d[1], rz = az - pr * d[2];
    return sqrtf(rx * rx + ry * ry + rz * rz);
}

static void co_bkernels(const CoEvolverPickupDef* d, const OString* ss, float* Bn, float* Bb)
{
    float ax = d->axis[0], ay = d->axis[1], az = d->axis[2];
    float al = sqrtf(ax * ax + ay * ay + az * az);
    if (al < 1e-9f) { ax = 0.0f; ay = 1.0f; az = 0.0f; }
    else { ax /= al; ay /= al; az /= al; }
    float i2s2 = (d->pole_sigma > 0.0f) ? 0.5f / (d->pole_sigma * d->pole_sigma) : 1e12f;
    float sp = d->coil_spacing;
    float au[3] = {ax, ay, az};
    for (int s = 0; s < ss->n_segs; ++s) {
        float m[3];
        for (int c = 0; c < 3; ++c) m[c] = 0.5f * (ss->xyz[s * 3 + c] + ss->xyz[(s + 1) * 3 + c]);
        float B;
        if (d->type == PICKUP_SINGLE_COIL) {
            float r = co_line_dist(m, d->pos, au);
            B = expf(-r * r * i2s2);
        } else if (d->type == PICKUP_HUMBUCKER) {
            float p1[3] = {d->pos[0] + 0.5f * sp * ax, d->pos[1] + 0.5f * sp * ay, d->pos[2] + 0.5f * sp * az};
            float p2[3] = {d->pos[0] - 0.5f * sp * ax, d->pos[1] - 0.5f * sp * ay, d->pos[2] - 0.5f * sp * az};
            float r1 = co_line_dist(m, p1, au), r2 = co_line_dist(m, p2, au);
            B = expf(-r1 * r1 * i2s2) - expf(-r2 * r2 * i2s2);
        } else {
            B = 0.0f;
        }
        const float* n = ss->nrm + s * 3;
        const float* b = ss->bin + s * 3;
        Bn[s] = B * (n[0] * ax + n[1] * ay + n[2] * az);
        Bb[s] = B * (b[0] * ax + b[1] * ay + b[2] * az);
    }
}

static int co_gidx(const OrcCoevo* st, int i, int j, int k)                   /* :786-793 */
{
    if (i < 0) i = 0; if (i >= st->Nx) i = st->Nx - 1;
    if (j < 0) j = 0; if (j >= st->Ny) j = st->Ny - 1;
    if (k < 0) k = 0; if (k >= st->Nz) k = st->Nz - 1;
    return k + st->Nz * (j + st->Ny * i);
}

static void co_bridge_kernel(OrcCoevo* st, int n, const float* xyz, float sigma)  /* :799-861 */
{
    int pz = st->plate_iz - 1; if (pz < 0) pz = 0;
    int rc = (int)ceilf(3.0f * sigma / st->dx) + 1;
    int mx = n * (2 * rc + 1) * (2 * rc + 1); if (mx < 1) mx = 1;
    int* idx = (int*)calloc(mx, sizeof(int));
    float* wgt = (float*)calloc(mx, sizeof(float));
    int cnt = 0;
    float i2s2 = 0.5f / (sigma * sigma);
    for (int s = 0; s < n; ++s) {
        float gx = (xyz[s * 3 + 0] - st->origin[0]) / st->dx;
        float gy = (xyz[s * 3 + 1] - st->origin[1]) / st->dx;
        int ci0 = (int)roundf(gx), cj0 = (int)roundf(gy);
        for (int di = -rc; di <= rc; ++di)
            for (int dj = -rc; dj <= rc; ++dj) {
                int ci = ci0 + di, cj = cj0 + dj;
                if (ci < 0 || ci >= st->Nx || cj < 0 || cj >= st->Ny) continue;
                float wx = (float)ci * st->dx + st->origin[0] - xyz[s * 3 + 0];
                float wy = (float)cj * st->dx + st->origin[1] - xyz[s * 3 + 1];
                float w = expf(-(wx * wx + wy * wy) * i2s2);
                if (w < 1e-4f) continue;
                int flat = co_gidx(st, ci, cj, pz);
                int dup = 0;
                for (int q = 0; q < cnt; ++q)
                    if (idx[q] == flat) { wgt[q] += w; dup = 1; break; }
                if (!dup && cnt < mx) { idx[cnt] = flat; wgt[cnt] = w; ++cnt; }
            }
    }
    float ws = 0.0f;
    for (int q = 0; q < cnt; ++q) ws += wgt[q];
    if (ws > 1e-12f) for (int q = 0; q < cnt; ++q) wgt[q] /= ws;
    st->bidx = idx; st->bwgt = wgt; st->nb = cnt;
}

static int co_neck_kernel(OrcCoevo* st, const uint8_t* pa, float sigma)        /* :868-940 */
{
    if (sigma <= 0.0f) sigma = 0.030f;
    float ymax = -FLT_MAX;
    for (int i = 0; i < st->Nx; ++i)
        for (int j = 0; j < st->Ny; ++j) {
            if (!pa[j + st->Ny * i]) continue;
            float y = (float)j * st->dx + st->origin[1];
            if (y > ymax) ymax = y;
        }
    if (ymax == -FLT_MAX) return -4;
    float cx = 0.0f, cy = ymax - 0.012f;
    int rc = (int)ceilf(3.0f * sigma / st->dx) + 1;
    int ci = (int)floorf((cx - st->origin[0]) / st->dx);
    int cj = (int)floorf((cy - st->origin[1]) / st->dx);
    int cap = (2 * rc + 1) * (2 * rc + 1);
    int* idx = (int*)calloc(cap, sizeof(int));
    float* wgt = (float*)calloc(cap, sizeof(float));
    float i2s2 = 0.5f / (sigma * sigma), ws = 0.0f;
    int cnt = 0;
    for (int di = -rc; di <= rc; ++di) {
        int i = ci + di;
        if (i < 0 || i >= st->Nx) continue;
        for (int dj = -rc; dj <= rc; ++dj) {
            int j = cj + dj;
            if (j < 0 || j >= st->Ny) continue;
            int p = j + st->Ny * i;
            if (!pa[p]) continue;
            float x = (float)i * st->dx + st->origin[0];
            float y = (float)j * st->dx + st->origin[1];
            float ddx = x - cx, ddy = y - cy;
            float ww = expf(-(ddx * ddx + ddy * ddy) * i2s2);
            if (ww <= 1e-12f) continue;
            idx[cnt] = p; wgt[cnt] = ww; ws += ww; ++cnt;
        }
    }
    if (cnt <= 0 || ws <= 1e-12f) { free(idx); free(wgt); return -4; }
    for (int q = 0; q < cnt; ++q) wgt[q] /= ws;
    st->nidx = idx; st->nwgt = wgt; st->nn = cnt;
    return 0;
}

static uint8_t* co_soundhole(const CoEvolverBodyDef* b)                        /* :946-981 */
{
    int Np = b->Nx * b->Ny;
    uint8_t* out = (uint8_t*)calloc(Np, 1);
    memcpy(out, b->plate_active, Np);
    const float cx = 0.0f, cy = 0.050f, r = 0.028f, r2 = r * r;
    float x0 = b->origin[0], y0 = b->origin[1];
    float x1 = x0 + (float)(b->Nx - 1) * b->dx, y1 = y0 + (float)(b->Ny - 1) * b->dx;
    if (cx + r < x0 || cx - r > x1 || cy + r < y0 || cy - r > y1) return out;
    for (int i = 0; i < b->Nx; ++i) {
        float x = b->origin[0] + (float)i * b->dx;
        for (int j = 0; j < b->Ny; ++j) {
            float y = b->origin[1] + (float)j * b->dx;
            float ddx = x - cx, ddy = y - cy;
            if (ddx * ddx + ddy * ddy < r2) out[j + b->Ny * i] = 0;
        }
    }
    return out;
}

SK_EXPORT void orc_coevo_destroy(OrcCoevo* st)
{
    if (!st) return;
    if (st->f) orc_destroy_f32(st->f);
    if (st->amr) amo_destroy(st->amr);
    if (st->s) for (int i = 0; i < st->ns; ++i) co_string_free(&st->s[i]);
    if (st->pk) for (int p = 0; p < st->npk; ++p) {
        if (st->pk[p].Bn) for (int i = 0; i < st->ns; ++i) free(st->pk[p].Bn[i]);
        if (st->pk[p].Bb) for (int i = 0; i < st->ns; ++i) free(st->pk[p].Bb[i]);
        free(st->pk[p].Bn); free(st->pk[p].Bb);
    }
    free(st->s); free(st->pk); free(st->mic);
    free(st->bidx); free(st->bwgt); free(st->nidx); free(st->nwgt);
    free(st->pk_out); free(st->mic_out); free(st->pk_acc);
    free(st->sad_idx); free(st->sad_wgt); free(st->sad_w);
    free(st->mP_i); free(st->mP_w);
    for (int a = 0; a < 3; ++a) { free(st->mV_i[a]); free(st->mV_w[a]); }
    free(st->sbk); free(st->bdrive); free(st->seg_off); free(st->mag_map);
    free(st->K0); free(st->K1); free(st->v0f); free(st->v1f); free(st->mag_tmp);
    free(st->ev);
    free(st);
}

/* Set-up shared by both bodies (:1063-1381, :1480-1745); the body, grid
 * fields, dt_fdtd, K and the bridge (and AMR neck) plate lists are in place. */
static int co_setup(OrcCoevo* st, int ns, const CoEvolverStringDef* sd, int npk,
                    const CoEvolverPickupDef* pd, int nmic, const CoEvolverMicDef* md,
                    const CoEvolverBodyDef* b, float bridge_sigma)
{
    st->ns = ns;
    st->s = (OString*)calloc(ns > 0 ? ns : 1, sizeof(OString));
    for (int i = 0; i < ns; ++i)
        if (co_string_init(&st->s[i], &sd[i], st->dt_fdtd)) return -1;
    for (int i = 0; i < ns; ++i) {
        if (sd[i].neck_freq_hz > 0.0f && sd[i].neck_mass_kg > 0.0f) {
            st->n_m = sd[i].neck_mass_kg;
            float om = 2.0f * (float)CO_PI * sd[i].neck_freq_hz;
            st->n_k = st->n_m * om * om;
            float Q = (sd[i].neck_Q > 0.0f) ? sd[i].neck_Q : 30.0f;
            st->n_c = st->n_m * om / Q;
            float oj = 2.0f * (float)CO_PI * 95.0f;
            st->nb_k = st->n_m * oj * oj;
            st->nb_c = st->n_m * oj / 2.0f;
            break;
        }
    }
    if (st->amr) {
        if (st->n_m > 0.0f && st->nn == 0) st->n_m = 0.0f;                 /* :1510-1512 */
    } else if (st->n_m > 0.0f && co_neck_kernel(st, b->plate_active, 0.030f)) {
        return -1;
    }
    st->npk = npk;
    st->pk = (OPickup*)calloc(npk > 0 ? npk : 1, sizeof(OPickup));
    for (int p = 0; p < npk; ++p) {
        OPickup* k = &st->pk[p];
        k->type = pd[p].type; k->mask = pd[p].string_mask; k->sens = pd[p].sensitivity;
        k->Bn = (float**)calloc(ns > 0 ? ns : 1, sizeof(float*));
        k->Bb = (float**)calloc(ns > 0 ? ns : 1, sizeof(float*));
        for (int i = 0; i < ns; ++i)
            if (pd[p].type != PICKUP_PIEZO) {
                k->Bn[i] = (float*)calloc(st->s[i].n_segs, sizeof(float));
                k->Bb[i] = (float*)calloc(st->s[i].n_segs, sizeof(float));
                co_bkernels(&pd[p], &st->s[i], k->Bn[i], k->Bb[i]);
            }
    }
    st->nmic = nmic;
    st->mic = (OMic*)calloc(nmic > 0 ? nmic : 1, sizeof(OMic));
    for (int m = 0; m < nmic; ++m) {
        for (int q = 0; q < 3; ++q) {
            st->mic[m].g[q] = (md[m].pos[q] - st->origin[q]) / st->dx;
            st->mic[m].axis[q] = md[m].axis[q];
        }
        st->mic[m].gain = md[m].gain; st->mic[m].pa = md[m].polar_a; st->mic[m].pb = md[m].polar_b;
        if (md[m].polar_b != 0.0f) st->any_vel = 1;
    }
    st->pk_out = (float*)calloc((size_t)(npk > 0 ? npk : 1) * CO_OUTPUT_BUF, sizeof(float));
    st->mic_out = (float*)calloc((size_t)(nmic > 0 ? nmic : 1) * CO_OUTPUT_BUF, sizeof(float));
    st->pk_acc = (float*)calloc(npk > 0 ? npk : 1, sizeof(float));
    if (ns > 0) {
        st->sad_idx = (int*)calloc(ns * 4, sizeof(int));
        st->sad_wgt = (float*)calloc(ns * 4, sizeof(float));
        st->sad_w = (float*)calloc(ns, sizeof(float));
        for (int i = 0; i < ns; ++i) {                                   /* :1160-1183 */
            const OString* ss = &st->s[i];
            int Nn = ss->n_nodes;
            float gi = (ss->xyz[(Nn - 1) * 3 + 0] - st->origin[0]) / st->dx;
            float gj = (ss->xyz[(Nn - 1) * 3 + 1] - st->origin[1]) / st->dx;
            if (gi < 0.0f) gi = 0.0f; if (gi > (float)(st->Nx - 1)) gi = (float)(st->Nx - 1);
            if (gj < 0.0f) gj = 0.0f; if (gj > (float)(st->Ny - 1)) gj = (float)(st->Ny - 1);
            int i0 = (int)gi, i1 = (i0 < st->Nx - 1) ? i0 + 1 : i0;
            int j0 = (int)gj, j1 = (j0 < st->Ny - 1) ? j0 + 1 : j0;
            float fi = gi - (float)i0, fj = gj - (float)j0;
            st->sad_idx[i * 4 + 0] = j0 + st->Ny * i0;
            st->sad_idx[i * 4 + 1] = j0 + st->Ny * i1;
            st->sad_idx[i * 4 + 2] = j1 + st->Ny * i0;
            st->sad_idx[i * 4 + 3] = j1 + st->Ny * i1;
            st->sad_wgt[i * 4 + 0] = (1.0f - fi) * (1.0f - fj);
            st->sad_wgt[i * 4 + 1] = fi * (1.0f - fj);
            st->sad_wgt[i * 4 + 2] = (1.0f - fi) * fj;
            st->sad_wgt[i * 4 + 3] = fi * fj;
        }
    }
    if (nmic > 0 && st->amr) {                                          /* :1613-1646 */
        float* pos = (float*)calloc(nmic * 3, sizeof(float));
        float* ax = (float*)calloc(nmic * 3, sizeof(float));
        for (int m = 0; m < nmic; ++m)
            for (int q = 0; q < 3; ++q) { pos[m * 3 + q] = md[m].pos[q]; ax[m * 3 + q] = md[m].axis[q]; }
        int rc = amo_setup_mics(st->amr, nmic, pos, ax, st->any_vel);
        free(pos); free(ax);
        if (rc != AMO_OK) return -1;
    } else if (nmic > 0) {                                               /* :1197-1246 */
        float* g = (float*)calloc(nmic * 3, sizeof(float));
        for (int m = 0; m < nmic; ++m) for (int q = 0; q < 3; ++q) g[m * 3 + q] = st->mic[m].g[q];
        st->mP_i = (int*)calloc(nmic * 8, sizeof(int));
        st->mP_w = (float*)calloc(nmic * 8, sizeof(float));
        orc_precompute_pressure_samplers_f32(st->f, nmic, g, st->mP_i, st->mP_w);
        if (st->any_vel) {
            for (int a = 0; a < 3; ++a) {
                st->mV_i[a] = (int*)calloc(nmic * 8, sizeof(int));
                st->mV_w[a] = (float*)calloc(nmic * 8, sizeof(float));
            }
            orc_precompute_velocity_samplers_f32(st->f, nmic, g, st->mV_i[0], st->mV_w[0],
                                                 st->mV_i[1], st->mV_w[1], st->mV_i[2],
                                                 st->mV_w[2]);
        }
        free(g);
    }
    st->seg_off = (int*)calloc(ns + 1, sizeof(int));                    /* :1254-1306 */
    for (int i = 0; i < ns; ++i) st->seg_off[i + 1] = st->seg_off[i] + st->s[i].n_segs;
    st->tot_seg = ns > 0 ? st->seg_off[ns] : 0;
    for (int p = 0; p < npk; ++p) if (pd[p].type != PICKUP_PIEZO) ++st->nmag;
    if (st->nmag > 0 && st->tot_seg > 0) {
        st->mag_map = (int*)calloc(st->nmag, sizeof(int));
        st->K0 = (float*)calloc((size_t)st->nmag * st->tot_seg, sizeof(float));
        st->K1 = (float*)calloc((size_t)st->nmag * st->tot_seg, sizeof(float));
        st->v0f = (float*)calloc(st->tot_seg, sizeof(float));
        st->v1f = (float*)calloc(st->tot_seg, sizeof(float));
        int r = 0;
        for (int p = 0; p < npk; ++p) {
            if (pd[p].type == PICKUP_PIEZO) continue;
            st->mag_map[r] = p;
            for (int i = 0; i < ns; ++i) {
                if (!(st->pk[p].mask & (1u << (unsigned)i))) continue;
                for (int s = 0; s < st->s[i].n_segs; ++s) {
                    st->K0[(size_t)r * st->tot_seg + st->seg_off[i] + s] = st->pk[p].Bn[i][s] * st->s[i].ds;
                    st->K1[(size_t)r * st->tot_seg + st->seg_off[i] + s] = st->pk[p].Bb[i][s] * st->s[i].ds;
                }
            }
            ++r;
        }
        st->mag_tmp = (float*)calloc(st->nmag, sizeof(float));
    }
    if (ns > 0 && st->nb > 0) {                                          /* :1322-1375 */
        int nb = st->nb;
        st->sbk = (float*)calloc((size_t)ns * nb, sizeof(float));
        st->bdrive = (float*)calloc(nb, sizeof(float));
        float sigma = bridge_sigma;
        float i2s2 = 0.5f / (sigma * sigma);
        for (int i = 0; i < ns; ++i) {
            const OString* ss = &st->s[i];
            int Nn = ss->n_nodes;
            float sx = ss->xyz[(Nn - 1) * 3 + 0], sy = ss->xyz[(Nn - 1) * 3 + 1];
            float* row = st->sbk + (size_t)i * nb;
            float ws = 0.0f;
            for (int c = 0; c < nb; ++c) {
                int flat = st->bidx[c];
                int t = flat / st->Nz;
                int j = t % st->Ny, ii = t / st->Ny;
                float cx = (float)ii * st->dx + st->origin[0];
                float cy = (float)j * st->dx + st->origin[1];
                float d2x = sx - cx, d2y = sy - cy;
                row[c] = expf(-(d2x * d2x + d2y * d2y) * i2s2);
                ws += row[c];
            }
            if (ws > 1e-12f) for (int c = 0; c < nb; ++c) row[c] /= ws;
        }
    }
    return 0;
}

/* coevolver_create — :987-1381 (the FDTD of fdtd_create: order 4, reference dt) */
SK_EXPORT OrcCoevo* orc_coevo_create(int ns, const CoEvolverStringDef* sd, int npk,
                                     const CoEvolverPickupDef* pd, int nmic,
                                     const CoEvolverMicDef* md, const CoEvolverBodyDef* b,
                                     float sample_rate, int modal_stride, float force_scale)
{
    (void)force_scale;
    if (!sd || !pd || !md || !b) return NULL;
    if (ns < 0 || npk < 0 || nmic < 0) return NULL;
    if (sample_rate <= 0 || modal_stride <= 0) return NULL;
    OrcCoevo* st = (OrcCoevo*)calloc(1, sizeof(OrcCoevo));
    uint8_t* pa = co_soundhole(b);
    OrcConfig cf;
    memset(&cf, 0, sizeof(cf));
    cf.Nx = b->Nx; cf.Ny = b->Ny; cf.Nz = b->Nz; cf.dx = b->dx; cf.c = b->c;
    cf.rho_air = b->rho_air; cf.plate_iz = b->plate_iz; cf.plate_mass = b->plate_mass_density;
    cf.plate_D = b->plate_stiffness_D; cf.plate_alpha_M = b->plate_alpha_M;
    cf.plate_beta_K = b->plate_beta_K; cf.n_pml = b->n_pml; cf.order = 4; cf.courant = 0.0f;
    cf.boundary = 0;
    for (int q = 0; q < 6; ++q) cf.refl[q] = 1.0f;
    st->f = orc_create_f32(&cf, b->cell_type, pa, NULL, NULL);
    free(pa);
    if (!st->f) goto fail;
    st->Nx = b->Nx; st->Ny = b->Ny; st->Nz = b->Nz; st->dx = b->dx;
    st->rho_air = b->rho_air; st->c = b->c; st->plate_iz = b->plate_iz;
    for (int q = 0; q < 3; ++q) st->origin[q] = b->origin[q];
    st->sr = sample_rate; st->dt_audio = 1.0f / sample_rate; st->dt_fdtd = orc_get_dt_f32(st->f);
    st->gamma_scale = 1.0f; st->modal_stride = modal_stride;
    st->K = (int)ceilf(st->dt_audio / st->dt_fdtd);
    if (st->K < 1) st->K = 1;
    if (st->K > CO_MAX_SUBSTEPS) st->K = CO_MAX_SUBSTEPS;
    if (b->n_bridge_src > 0 && b->bridge_src_xyz)
        co_bridge_kernel(st, b->n_bridge_src, b->bridge_src_xyz, b->bridge_sigma);
    if (st->nb > 0) orc_set_bridge_sources_f32(st->f, st->nb, st->bidx, st->bwgt);
    if (co_setup(st, ns, sd, npk, pd, nmic, md, b,
                 (b->bridge_sigma > 0.0f) ? b->bridge_sigma : st->dx)) goto fail;
    return st;
fail:
    orc_coevo_destroy(st);
    return NULL;
}

/* coevolver_create_amr — :1387-1746: the AMR pressure backend of
 * amr_pressure_backend_create (acoustic_pressure_backend.cpp:273-569:
 * create, plate, bridge and optional neck sources, border condition) and
 * the plate grid in place of the uniform grid.  *stage gets the failure
 * label. */
SK_EXPORT OrcCoevo* orc_coevo_create_amr(int ns, const CoEvolverStringDef* sd, int npk,
                                         const CoEvolverPickupDef* pd, int nmic,
                                         const CoEvolverMicDef* md,
                                         const AMRCoevolverDescriptor* d, float sample_rate,
                                         int modal_stride, const char** stage)
{
    *stage = "";
    if (ns < 0 || npk < 0 || nmic < 0) { *stage = "invalid_counts"; return NULL; }
    if ((ns > 0 && !sd) || (npk > 0 && !pd) || (nmic > 0 && !md) || !d) {
        *stage = "invalid_null_args"; return NULL;
    }
    if (sample_rate <= 0 || modal_stride <= 0) { *stage = "invalid_timing_args"; return NULL; }
    OrcCoevo* st = (OrcCoevo*)calloc(1, sizeof(OrcCoevo));
    *stage = "pressure_backend_create";
    if (d->n_cells <= 0 || d->n_faces <= 0 || d->n_bridge_plate <= 0 || !d->bridge_plate_idx ||
        !d->bridge_plate_wgt || d->plate_Nx <= 0 || d->plate_Ny <= 0 || d->plate_dx <= 0.0f ||
        d->plate_mass_density <= 0.0f || d->plate_stiffness_D <= 0.0f) goto fail;
    st->amr = amo_create(d->n_cells, d->cell_centers, d->cell_volumes, d->open_volume_frac,
                         d->cell_types, d->n_faces, d->face_cell_neg, d->face_cell_pos,
                         d->face_area, d->face_open_frac, d->face_distance, d->c, d->rho_air,
                         d->min_dx, d->gradient_order);
    if (!st->amr) goto fail;
    if (amo_setup_plate(st->amr, d->plate_Nx, d->plate_Ny, d->plate_dx, d->plate_origin,
                        d->plate_active, d->plate_mass_density, d->plate_stiffness_D,
                        d->plate_alpha_M, d->plate_beta_K, d->n_active_plate,
                        d->plate_active_flat_idx, d->plate_face_above_starts,
                        d->plate_face_above_idx, d->plate_face_above_wgt,
                        d->plate_face_below_starts, d->plate_face_below_idx,
                        d->plate_face_below_wgt, d->plate_cell_above, d->plate_cell_below) != AMO_OK)
        goto fail;
    if (amo_set_bridge_sources(st->amr, d->n_bridge_plate, d->bridge_plate_idx,
                               d->bridge_plate_wgt) != AMO_OK) goto fail;
    if (d->n_neck_plate > 0 && d->neck_plate_idx && d->neck_plate_wgt &&
        amo_set_neck_sources(st->amr, d->n_neck_plate, d->neck_plate_idx,
                             d->neck_plate_wgt) != AMO_OK) goto fail;
    if (d->n_pml > 0 &&
        amo_set_border_condition(st->amr, d->border_mode,
                                 d->border_sigma_order > 0.0f ? d->border_sigma_order : 3.0f,
                                 d->border_R_reflection, d->border_Z_match, d->n_pml,
                                 d->bounds_min, d->bounds_max) != AMO_OK) goto fail;
    st->Nx = d->plate_Nx; st->Ny = d->plate_Ny; st->Nz = 1; st->dx = d->plate_dx;
    for (int q = 0; q < 3; ++q) st->origin[q] = d->plate_origin[q];
    st->rho_air = (float)d->rho_air; st->c = (float)d->c;
    st->sr = sample_rate; st->dt_audio = 1.0f / sample_rate;
    st->dt_fdtd = (float)amo_get_dt(st->amr);
    *stage = "dt_fdtd_invalid";
    if (st->dt_fdtd <= 0.0f) goto fail;
    st->gamma_scale = 1.0f; st->modal_stride = modal_stride;
    st->K = (int)ceilf(st->dt_audio / st->dt_fdtd);
    if (st->K < 1) st->K = 1;
    if (st->K > CO_MAX_SUBSTEPS) st->K = CO_MAX_SUBSTEPS;
    st->nb = d->n_bridge_plate;
    st->bidx = (int*)malloc(st->nb * sizeof(int));
    st->bwgt = (float*)malloc(st->nb * sizeof(float));
    memcpy(st->bidx, d->bridge_plate_idx, st->nb * sizeof(int));
    memcpy(st->bwgt, d->bridge_plate_wgt, st->nb * sizeof(float));
    if (d->n_neck_plate > 0 && d->neck_plate_idx && d->neck_plate_wgt) {
        st->nn = d->n_neck_plate;
        st->nidx = (int*)malloc(st->nn * sizeof(int));
        st->nwgt = (float*)malloc(st->nn * sizeof(float));
        memcpy(st->nidx, d->neck_plate_idx, st->nn * sizeof(int));
        memcpy(st->nwgt, d->neck_plate_wgt, st->nn * sizeof(float));
    }
    *stage = "setup";
    if (co_setup(st, ns, sd, npk, pd, nmic, md, NULL,
                 (d->plate_dx > 0.0f) ? 2.0f * d->plate_dx : 0.005f)) goto fail;
    *stage = "";
    return st;
fail:
    orc_coevo_destroy(st);
    return NULL;
}

/* ── string sub-step pieces ───────────────────────────────────────────── */

static int co_validate(const OString* ss)                                      /* :538-561 */
{
    float arc = 0.0f;
    for (int i = 0; i < ss->n_nodes - 1; ++i) {
        float d0 = ss->u[0][i + 1] - ss->u[0][i];
        float d1 = ss->u[1][i + 1] - ss->u[1][i];
        if (!isfinite(d0) || !isfinite(d1)) return -3;
        float sl = sqrtf(d0 * d0 + d1 * d1) / ss->ds;
        if (!isfinite(sl) || sl > CO_MAX_SLOPE) return -3;
        arc += sqrtf(ss->ds * ss->ds + d0 * d0 + d1 * d1);
    }
    float fr = (arc - ss->L) / ss->L;
    if (!isfinite(fr) || fr > CO_MAX_STRETCH) return -3;
    return 0;
}

static int co_tension(OString* ss)                                             /* :563-593 */
{
    float ds2 = ss->ds * ss->ds, arc = 0.0f;
    for (int i = 0; i < ss->n_nodes - 1; ++i) {
        float d0 = ss->u[0][i + 1] - ss->u[0][i];
        float d1 = ss->u[1][i + 1] - ss->u[1][i];
        if (!isfinite(d0) || !isfinite(d1)) return -3;
        arc += sqrtf(ds2 + d0 * d0 + d1 * d1);
    }
    float strain = (arc - ss->L) / ss->L;
    if (!isfinite(strain)) return -3;
    if (strain < 0.0f) strain = 0.0f;
    ss->T_eff = ss->T0 + ss->EA * strain;
    if (!isfinite(ss->T_eff) || ss->T_eff <= 0.0f) return -3;
    float cfl = sqrtf(ss->T_eff / ss->mu) * ss->dt_s / ss->ds;
    if (!isfinite(cfl) || cfl * cfl > 0.9025f) return -3;
    return 0;
}

static int co_step_pol(OString* ss, int pol, float nut)                         /* :595-663 */
{
    int N = ss->n_nodes;
    float cfl = sqrtf(ss->T_eff / ss->mu) * ss->dt_s / ss->ds;
    float c2 = cfl * cfl;
    float h = 0.5f * ss->gamma_eff * ss->dt_s;
    float inv = 1.0f / (1.0f + h);
    float omh = 1.0f - h;
    float *u = ss->u[pol], *up = ss->up[pol], *ut = ss->ut[pol], *fe = ss->fe[pol];
    ut[N - 1] = (pol == 0) ? ss->bridge_w : 0.0f;
    ut[0] = (pol == 0) ? nut : 0.0f;
    if (N > 2) {
        for (int i = 1; i <= N - 2; ++i)
            ut[i] = 2.0f * u[i] - omh * up[i] + c2 * (u[i + 1] - 2.0f * u[i] + u[i - 1]) + fe[i];
        if (N > 4 && ss->biharm > 0.0f)
            for (int i = 2; i <= N - 3; ++i)
                ut[i] -= ss->biharm * (u[i - 2] - 4.0f * u[i - 1] + 6.0f * u[i] - 4.0f * u[i + 1] + u[i + 2]);
        int bad = 0;
        for (int i = 1; i <= N - 2; ++i) {
            ut[i] *= inv;
            if (!isfinite(ut[i])) bad = 1;
        }
        if (bad) return -3;
    }
    float* t = ss->up[pol]; ss->up[pol] = ss->u[pol]; ss->u[pol] = ss->ut[pol]; ss->ut[pol] = t;
    memset(fe, 0, N * sizeof(float));
    return 0;
}

static void co_modal(OString* ss)                                              /* :665-689 */
{
    int N = ss->n_nodes;
    float sc = 2.0f / (float)(N > 1 ? N - 1 : 1);
    for (int k = 0; k < CO_N_MODES; ++k) {
        float a = 0.0f, b = 0.0f;
        for (int i = 0; i < N; ++i) {
            a += ss->basis[k * N + i] * ss->u[0][i];
            b += ss->basis[k * N + i] * ss->up[0][i];
        }
        ss->modal_re[k] = a * sc;
        ss->modal_im[k] = b * sc;
    }
}

/* _step_one_sample — :1910-2226 */
static int co_sample(OrcCoevo* st, int sample_abs)
{
    int K = st->K, nb = st->nb;
    for (int p = 0; p < st->npk; ++p) st->pk_acc[p] = 0.0f;
    float per_sub[64];
    memset(per_sub, 0, sizeof(per_sub));
    for (int i = 0; i < st->ns && i < 64; ++i) {
        OString* ss = &st->s[i];
        ss->gamma_eff = ss->gamma * st->gamma_scale;
        int bi = sample_abs - ss->fonset;
        float ft = (bi >= 0 && bi < ss->nfbuf) ? ss->fbuf[bi] : 0.0f;
        int tot = K * ss->n_sub;
        per_sub[i] = (tot > 0 && ft != 0.0f) ? ft / (float)tot : 0.0f;
    }
    for (int sub = 0; sub < K; ++sub) {
        float neck_force = 0.0f;
        if (st->n_m > 0.0f && st->nn > 0) {                              /* :1934-1983 */
            float pw;
            if (st->amr) amo_sample_plate_weighted(st->amr, st->nn, st->nidx, st->nwgt, &pw);
            else orc_sample_plate_weighted_f32(st->f, st->nn, st->nidx, st->nwgt, &pw);
            float fs = 0.0f;
            for (int i = 0; i < st->ns; ++i) {
                OString* ss = &st->s[i];
                if (ss->n_nodes > 1) fs += ss->T_eff * (ss->u[0][1] - st->neck_w) / ss->ds;
            }
            float dt_n = st->dt_fdtd; if (dt_n <= 0.0f) dt_n = st->dt_audio;
            float idt = 1.0f / dt_n, idt2 = idt * idt;
            float pv = (pw - st->neck_plate_w_prev) * idt;
            float m = st->n_m, c = st->n_c + st->nb_c, k = st->n_k + st->nb_k;
            float den = m * idt2 + 0.5f * c * idt;
            float rhs = fs + st->nb_k * pw + st->nb_c * pv + (2.0f * m * idt2 - k) * st->neck_w
                      + (-m * idt2 + 0.5f * c * idt) * st->neck_w_prev;
            float nw = (den > 1e-12f) ? rhs / den : pw;
            if (!isfinite(nw) || fabsf(nw) > 0.05f) return -3;
            st->neck_w_prev = st->neck_w;
            st->neck_w = nw;
            float nv = (st->neck_w - st->neck_w_prev) * idt;
            neck_force = st->nb_k * (st->neck_w - pw) + st->nb_c * (nv - pv);
            if (!isfinite(neck_force)) return -3;
            st->neck_plate_w_prev = pw;
        }
        for (int i = 0; i < st->ns; ++i) {                               /* :1987-2003 */
            OString* ss = &st->s[i];
            for (int q = 0; q < ss->n_sub; ++q) {
                if (co_tension(ss)) return -3;
                if (per_sub[i] != 0.0f) ss->fe[0][ss->fnode] += per_sub[i];
                float nut = (st->n_m > 0.0f) ? st->neck_w : 0.0f;
                if (co_step_pol(ss, 0, nut)) return -3;
                if (co_step_pol(ss, 1, 0.0f)) return -3;
                if (co_validate(ss)) return -3;
            }
        }
        if (nb > 0 && st->bdrive && st->sbk) {                           /* :2009-2062 */
            float fpl[64];
            float dt_b = st->dt_fdtd; if (dt_b <= 0.0f) dt_b = st->dt_audio;
            int nf = st->ns < 64 ? st->ns : 64;
            for (int i = 0; i < nf; ++i) {
                OString* ss = &st->s[i];
                int Nn = ss->n_nodes;
                float sl = (ss->u[0][Nn - 2] - ss->bridge_w) / ss->ds;
                if (!isfinite(sl) || fabsf(sl) > 1.0f) return -3;
                float fstr = ss->T_eff * sl;
                float m = ss->b_m, c = ss->b_c, k = ss->b_k;
                float idt = 1.0f / dt_b, idt2 = idt * idt;
                float pv = (ss->saddle_w - ss->saddle_w_prev) * idt;
                float den = m * idt2 + 0.5f * c * idt;
                float rhs = fstr + k * ss->saddle_w + c * pv + (2.0f * m * idt2 - k) * ss->bridge_w
                          + (-m * idt2 + 0.5f * c * idt) * ss->bridge_w_prev;
                float bn = (den > 1e-12f) ? rhs / den : ss->saddle_w;
                if (!isfinite(bn) || fabsf(bn) > 0.05f) return -3;
                ss->bridge_w_prev = ss->bridge_w;
                ss->bridge_w = bn;
                float bv = (ss->bridge_w - ss->bridge_w_prev) * idt;
                float fp = k * (ss->bridge_w - ss->saddle_w) + c * (bv - pv);
                if (!isfinite(fp)) return -3;
                fpl[i] = fp;
                ss->saddle_w_prev = ss->saddle_w;
            }
            for (int c = 0; c < nb; ++c) {
                float a = 0.0f;
                for (int i = 0; i < nf; ++i) a += st->sbk[(size_t)i * nb + c] * fpl[i];
                st->bdrive[c] = a;
            }
        }
        if (st->amr) {                                                    /* :2064-2119 */
            if (nb > 0 && st->bdrive) amo_inject_bridge_drive(st->amr, st->bdrive, nb, 1.0f);
            if (neck_force != 0.0f && st->nn > 0)
                amo_inject_plate_force(st->amr, st->nn, st->nidx, st->nwgt, neck_force);
            if (amo_step(st->amr, 1)) return -3;
            amo_sample_plate_batch(st->amr, st->ns, st->sad_idx, st->sad_wgt, st->sad_w);
        } else {
            if (nb > 0 && st->bdrive) orc_inject_bridge_drive_f32(st->f, st->bdrive, 1.0f);
            if (neck_force != 0.0f && st->nn > 0)
                orc_inject_plate_force_f32(st->f, st->nn, st->nidx, st->nwgt, neck_force);
            int rc = orc_step_f32(st->f, 1);
            if (rc) return -3;
            orc_sample_plate_batch_f32(st->f, st->ns, st->sad_idx, st->sad_wgt, st->sad_w);
        }
        for (int i = 0; i < st->ns; ++i) st->s[i].saddle_w = st->sad_w[i];
        ++st->modal_counter;
        if (st->modal_counter >= st->modal_stride) {
            st->modal_counter = 0;
            for (int i = 0; i < st->ns; ++i) co_modal(&st->s[i]);
        }
    }
    if (st->nmag > 0 && st->tot_seg > 0) {                               /* :2135-2171 */
        for (int i = 0; i < st->ns; ++i) {
            const OString* ss = &st->s[i];
            float idt = 1.0f / ss->dt_s;
            for (int s = 0; s < ss->n_segs; ++s) {
                st->v0f[st->seg_off[i] + s] = (ss->u[0][s] - ss->up[0][s]) * idt;
                st->v1f[st->seg_off[i] + s] = (ss->u[1][s] - ss->up[1][s]) * idt;
            }
        }
        for (int r = 0; r < st->nmag; ++r) {
            const float* k0 = st->K0 + (size_t)r * st->tot_seg;
            const float* k1 = st->K1 + (size_t)r * st->tot_seg;
            float a = 0.0f, b = 0.0f;
            for (int j = 0; j < st->tot_seg; ++j) a += k0[j] * st->v0f[j];
            for (int j = 0; j < st->tot_seg; ++j) b += k1[j] * st->v1f[j];
            st->mag_tmp[r] = a + b;
            st->pk_acc[st->mag_map[r]] = st->mag_tmp[r];
        }
    }
    for (int p = 0; p < st->npk; ++p) {                                  /* :2174-2181, :1892-1904 */
        OPickup* k = &st->pk[p];
        if (k->type != PICKUP_PIEZO) continue;
        float sig = 0.0f;
        for (int i = 0; i < st->ns; ++i) {
            float v = 0.0f;
            if (k->mask & (1u << (unsigned)i)) {
                const OString* ss = &st->s[i];
                int N = ss->n_nodes;
                float s0 = (ss->u[0][N - 1] - ss->u[0][N - 2]) / ss->ds;
                float s1 = (ss->u[1][N - 1] - ss->u[1][N - 2]) / ss->ds;
                v = k->sens * ss->T_eff * sqrtf(s0 * s0 + s1 * s1);
            }
            sig += v;
        }
        st->pk_acc[p] = sig;
    }
    int op = st->out_pos;
    if (st->nmic > 0 && st->amr) {                                       /* :2185-2219 */
        float* mv = (float*)calloc((size_t)4 * st->nmic, sizeof(float));
        amo_sample_mics(st->amr, st->nmic, mv, mv + st->nmic, mv + 2 * st->nmic, mv + 3 * st->nmic);
        for (int m = 0; m < st->nmic; ++m) {
            const OMic* ms = &st->mic[m];
            float sig = mv[m] * ms->pa;
            if (ms->pb != 0.0f) {
                float vn = mv[st->nmic + m] * ms->axis[0] + mv[2 * st->nmic + m] * ms->axis[1]
                         + mv[3 * st->nmic + m] * ms->axis[2];
                sig += ms->pb * (st->rho_air * st->c) * vn;
            }
            st->mic_out[(size_t)m * CO_OUTPUT_BUF + op] = sig * ms->gain;
        }
        free(mv);
    } else if (st->nmic > 0) {                                           /* :2183-2219 */
        const OrcState_f32* f = st->f;
        for (int m = 0; m < st->nmic; ++m) {
            const int* id = st->mP_i + 8 * m;
            const float* w = st->mP_w + 8 * m;
            float P = w[0] * f->P[id[0]] + w[1] * f->P[id[1]] + w[2] * f->P[id[2]] + w[3] * f->P[id[3]]
                    + w[4] * f->P[id[4]] + w[5] * f->P[id[5]] + w[6] * f->P[id[6]] + w[7] * f->P[id[7]];
            float V[3] = {0.0f, 0.0f, 0.0f};
            if (st->any_vel) {
                const float* F[3] = {f->vx, f->vy, f->vz};
                for (int a = 0; a < 3; ++a) {
                    float v = 0.0f;
                    for (int t = 0; t < 8; ++t) v += st->mV_w[a][8 * m + t] * F[a][st->mV_i[a][8 * m + t]];
                    V[a] = v;
                }
            }
            const OMic* ms = &st->mic[m];
            float sig = P * ms->pa;
            if (ms->pb != 0.0f) {
                float vn = V[0] * ms->axis[0] + V[1] * ms->axis[1] + V[2] * ms->axis[2];
                sig += ms->pb * (st->rho_air * st->c) * vn;
            }
            st->mic_out[(size_t)m * CO_OUTPUT_BUF + op] = sig * ms->gain;
        }
    }
    for (int p = 0; p < st->npk; ++p) st->pk_out[(size_t)p * CO_OUTPUT_BUF + op] = st->pk_acc[p];
    st->out_pos = (op + 1) % CO_OUTPUT_BUF;
    ++st->written;
    return 0;
}

SK_EXPORT int orc_coevo_pluck(OrcCoevo* st, int si, float pos_n, float amp_in, int ndur)  /* :2413-2470 */
{
    if (si < 0 || si >= st->ns) return CE_ERR_STRING_IDX;
    OString* ss = &st->s[si];
    int N = ss->n_nodes;
    float pos = fmaxf(1e-3f, fminf(1.0f - 1e-3f, pos_n));
    float amp = fmaxf(-0.006f, fminf(0.006f, amp_in));
    int ip = (int)roundf(pos * (float)(N - 1));
    if (ip <= 0) ip = 1;
    if (ip >= N - 1) ip = N - 2;
    float ll = fmaxf((float)ip * ss->ds, ss->ds);
    float rl = fmaxf((float)(N - 1 - ip) * ss->ds, ss->ds);
    float F = ss->T0 * fabsf(amp) * (1.0f / ll + 1.0f / rl);
    if (amp < 0.0f) F = -F;
    int nd = (ndur > 0) ? ndur : (int)roundf(0.020f * st->sr);
    if (nd < 4) nd = 4;
    float sc = ss->dt_s * ss->dt_s / (ss->mu * ss->ds);
    free(ss->fbuf);
    ss->fbuf = (float*)malloc(nd * sizeof(float));
    for (int s = 0; s < nd; ++s) {
        float t = (float)s / (float)(nd - 1);
        float env = 0.5f * (1.0f - cosf((float)CO_PI * t));
        ss->fbuf[s] = F * env * sc;
    }
    ss->nfbuf = nd; ss->fnode = ip; ss->fonset = st->cur_sample;
    return 0;
}

SK_EXPORT int orc_coevo_inject_force(OrcCoevo* st, int si, int n, const float* f, float pos_n)  /* :2472-2509 */
{
    if (si < 0 || si >= st->ns) return CE_ERR_STRING_IDX;
    if (n <= 0) return 0;
    OString* ss = &st->s[si];
    int N = ss->n_nodes;
    float pos = fmaxf(0.0f, fminf(1.0f, pos_n));
    int i0 = (int)roundf(pos * (float)(N - 1));
    if (i0 <= 0) i0 = 1;
    if (i0 >= N - 1) i0 = N - 2;
    free(ss->fbuf);
    ss->fbuf = (float*)malloc(n * sizeof(float));
    float sc = ss->dt_s * ss->dt_s / (ss->mu * ss->ds);
    for (int s = 0; s < n; ++s) ss->fbuf[s] = f[s] * sc;
    ss->nfbuf = n; ss->fnode = i0; ss->fonset = st->written;
    return 0;
}

SK_EXPORT int orc_coevo_schedule_pluck(OrcCoevo* st, int onset, int si, float pos, float amp,
                                       int ndur)                                      /* :1756-1786 */
{
    if (si < 0 || si >= st->ns) return CE_ERR_PARAM;
    if (onset < 0) onset = 0;
    if (st->nev >= st->cap_ev) {
        st->cap_ev = st->cap_ev < 8 ? 8 : st->cap_ev * 2;
        st->ev = (OPluck*)realloc(st->ev, st->cap_ev * sizeof(OPluck));
    }
    int q = st->nev++;
    while (q > 0 && st->ev[q - 1].onset > onset) { st->ev[q] = st->ev[q - 1]; --q; }
    st->ev[q].onset = onset; st->ev[q].si = si; st->ev[q].pos = pos; st->ev[q].amp = amp;
    st->ev[q].ndur = ndur;
    return 0;
}

SK_EXPORT void orc_coevo_clear_schedule(OrcCoevo* st) { st->nev = 0; st->cursor = 0; }

static void co_fire(OrcCoevo* st, int sample_idx)
{
    while (st->cursor < st->nev && st->ev[st->cursor].onset <= sample_idx) {
        OPluck* e = &st->ev[st->cursor++];
        orc_coevo_pluck(st, e->si, e->pos, e->amp, e->ndur);
    }
}

/* coevolver_step — :2229-2251 */
SK_EXPORT int orc_coevo_step(OrcCoevo* st, int n)
{
    if (n <= 0) return 0;
    int base = st->written;
    for (int s = 0; s < n; ++s) {
        st->cur_sample = base + s;
        co_fire(st, base + s);
        int rc = co_sample(st, base + s);
        if (rc) return rc;
    }
    return 0;
}

/* The async worker's loop — :2323-2353 (block-relative sample index; the
 * pluck onset keeps the last synchronous sample index). */
SK_EXPORT int orc_coevo_step_worker(OrcCoevo* st, int n)
{
    int base = st->written;
    for (int s = 0; s < n; ++s) {
        co_fire(st, base + s);
        int rc = co_sample(st, s);
        if (rc) return rc;
    }
    return 0;
}

/* coevolver_step_block_with_drive — :2253-2321 */
SK_EXPORT int orc_coevo_step_block_with_drive(OrcCoevo* st, const float* const* blocks, int ns,
                                              int n, float pos_n, float* mic0)
{
    if (!mic0) return CE_ERR_NULL;
    if (ns != st->ns) return CE_ERR_PARAM;
    if (n < 0) return CE_ERR_DIM;
    if (n == 0) return 0;
    if (st->nmic < 1) return CE_ERR_DIM;
    int node[64]; float sc[64];
    for (int i = 0; i < ns; ++i) {
        OString* ss = &st->s[i];
        int N = ss->n_nodes;
        float pos = fmaxf(0.0f, fminf(1.0f, pos_n));
        int i0 = (int)roundf(pos * (float)(N - 1));
        if (i0 <= 0) i0 = 1;
        if (i0 >= N - 1) i0 = N - 2;
        node[i] = i0;
        sc[i] = ss->dt_s * ss->dt_s / (ss->mu * ss->ds);
    }
    int base = st->written;
    for (int s = 0; s < n; ++s) {
        co_fire(st, base + s);
        if (blocks)
            for (int i = 0; i < ns; ++i) {
                if (!blocks[i]) continue;
                float f = blocks[i][s];
                if (f != 0.0f) st->s[i].fe[0][node[i]] += f * sc[i];
            }
        int rc = co_sample(st, -1);
        if (rc) return rc;
        int last = (st->out_pos - 1 + CO_OUTPUT_BUF) % CO_OUTPUT_BUF;
        mic0[s] = st->mic_out[last];
    }
    return 0;
}

/* coevolver_reset — :2698-2738 */
SK_EXPORT int orc_coevo_reset(OrcCoevo* st)
{
    if (st->f) orc_reset_f32(st->f);          /* an AMR body is not reset (:2702) */
    for (int i = 0; i < st->ns; ++i) {
        OString* ss = &st->s[i];
        for (int p = 0; p < 2; ++p) {
            memset(ss->u[p], 0, ss->n_nodes * sizeof(float));
            memset(ss->up[p], 0, ss->n_nodes * sizeof(float));
            memset(ss->fe[p], 0, ss->n_nodes * sizeof(float));
        }
        memset(ss->modal_re, 0, sizeof(ss->modal_re));
        memset(ss->modal_im, 0, sizeof(ss->modal_im));
        ss->saddle_w = ss->saddle_w_prev = ss->bridge_w = ss->bridge_w_prev = 0.0f;
        ss->gamma_eff = ss->gamma;
        ss->nfbuf = 0;
    }
    st->gamma_scale = 1.0f;
    memset(st->pk_out, 0, (size_t)(st->npk > 0 ? st->npk : 1) * CO_OUTPUT_BUF * sizeof(float));
    memset(st->mic_out, 0, (size_t)(st->nmic > 0 ? st->nmic : 1) * CO_OUTPUT_BUF * sizeof(float));
    st->out_pos = 0; st->written = 0;
    st->neck_w = st->neck_w_prev = st->neck_plate_w_prev = 0.0f;
    st->modal_counter = 0; st->cursor = 0;
    return 0;
}

SK_EXPORT void orc_coevo_set_damping_scale(OrcCoevo* st, float s)             /* :2752-2756 */
{
    st->gamma_scale = (s > 0.0f) ? s : 1.0f;
}

SK_EXPORT int orc_coevo_set_face_fractions(OrcCoevo* st, const float* fx, const float* fy,
                                           const float* fz)
{
    return orc_set_face_fractions_f32(st->f, fx, fy, fz);
}

static int co_ring(const float* ring, int wp, float* out, int n)               /* :2511-2518 */
{
    if (n > CO_OUTPUT_BUF) return CE_ERR_DIM;
    int start = (wp - n + CO_OUTPUT_BUF) % CO_OUTPUT_BUF;
    for (int i = 0; i < n; ++i) out[i] = ring[(start + i) % CO_OUTPUT_BUF];
    return 0;
}

SK_EXPORT int orc_coevo_get_pickup_output(const OrcCoevo* st, int p, float* out, int n)
{
    if (p < 0 || p >= st->npk) return CE_ERR_DIM;
    return co_ring(st->pk_out + (size_t)p * CO_OUTPUT_BUF, st->out_pos, out, n);
}

SK_EXPORT int orc_coevo_get_mic_output(const OrcCoevo* st, int m, float* out, int n)
{
    if (m < 0 || m >= st->nmic) return CE_ERR_DIM;
    return co_ring(st->mic_out + (size_t)m * CO_OUTPUT_BUF, st->out_pos, out, n);
}

SK_EXPORT OrcState_f32* orc_coevo_fdtd(OrcCoevo* st) { return st->f; }
SK_EXPORT Amo* orc_coevo_amr(OrcCoevo* st) { return st->amr; }
SK_EXPORT int orc_coevo_substeps(const OrcCoevo* st) { return st->K; }
SK_EXPORT int orc_coevo_written(const OrcCoevo* st) { return st->written; }
SK_EXPORT float orc_coevo_dt_fdtd(const OrcCoevo* st) { return st->dt_fdtd; }
SK_EXPORT int orc_coevo_n_bridge(const OrcCoevo* st) { return st->nb; }
SK_EXPORT int orc_coevo_n_neck(const OrcCoevo* st) { return st->nn; }

/* Full string state of one string: u[0], u[1], up[0], up[1] (n_nodes each)
 * then T_eff, saddle_w, saddle_w_prev, bridge_w, bridge_w_prev, n_sub. */
SK_EXPORT int orc_coevo_string_state(const OrcCoevo* st, int si, float* out, int len)
{
    if (si < 0 || si >= st->ns) return CE_ERR_STRING_IDX;
    const OString* ss = &st->s[si];
    int N = ss->n_nodes;
    if (len != 4 * N + 6) return CE_ERR_DIM;
    memcpy(out, ss->u[0], N * sizeof(float));
    memcpy(out + N, ss->u[1], N * sizeof(float));
    memcpy(out + 2 * N, ss->up[0], N * sizeof(float));
    memcpy(out + 3 * N, ss->up[1], N * sizeof(float));
    out[4 * N + 0] = ss->T_eff;
    out[4 * N + 1] = ss->saddle_w;
    out[4 * N + 2] = ss->saddle_w_prev;
    out[4 * N + 3] = ss->bridge_w;
    out[4 * N + 4] = ss->bridge_w_prev;
    out[4 * N + 5] = (float)ss->n_sub;
    return 0;
}

/* neck_w, neck_w_prev, neck_plate_w_prev */
SK_EXPORT void orc_coevo_neck_state(const OrcCoevo* st, float* out)
{
    out[0] = st->neck_w; out[1] = st->neck_w_prev; out[2] = st->neck_plate_w_prev;
}

SK_EXPORT int orc_coevo_modal(const OrcCoevo* st, int si, float* re, float* im, int n)
{
    if (si < 0 || si >= st->ns) return CE_ERR_STRING_IDX;
    int m = n < CO_N_MODES ? n : CO_N_MODES;
    for (int k = 0; k < m; ++k) { re[k] = st->s[si].modal_re[k]; im[k] = st->s[si].modal_im[k]; }
    return 0;
}

/* String frames and parameters: tan, nrm, bin (n_segs*3 each), then ds, L, dt_s, biharm. */
SK_EXPORT int orc_coevo_string_frame(const OrcCoevo* st, int si, float* out, int len)
{
    if (si < 0 || si >= st->ns) return CE_ERR_STRING_IDX;
    const OString* ss = &st->s[si];
    int S = ss->n_segs;
    if (len != 9 * S + 4) return CE_ERR_DIM;
    memcpy(out, ss->tan, 3 * S * sizeof(float));
    memcpy(out + 3 * S, ss->nrm, 3 * S * sizeof(float));
    memcpy(out + 6 * S, ss->bin, 3 * S * sizeof(float));
    out[9 * S + 0] = ss->ds; out[9 * S + 1] = ss->L; out[9 * S + 2] = ss->dt_s;
    out[9 * S + 3] = ss->biharm;
    return 0;
}

/* Test hook: overwrite one string's displacement fields (u[0], u[1],
 * u_prev[0], u_prev[1]; NULL keeps a field). */
SK_EXPORT int orc_coevo_set_string(OrcCoevo* st, int si, const float* u0, const float* u1,
                                   const float* up0, const float* up1)
{
    if (si < 0 || si >= st->ns) return CE_ERR_STRING_IDX;
    OString* ss = &st->s[si];
    size_t b = ss->n_nodes * sizeof(float);
    if (u0) memcpy(ss->u[0], u0, b);
    if (u1) memcpy(ss->u[1], u1, b);
    if (up0) memcpy(ss->up[0], up0, b);
    if (up1) memcpy(ss->up[1], up1, b);
    return 0;
}

/* Sub-steps per body step of one string and the string's ds, dt_s, c2 at T0. */
SK_EXPORT void orc_coevo_string_params(const OrcCoevo* st, int si, float* out4)
{
    const OString* ss = &st->s[si];
    out4[0] = (float)ss->n_sub; out4[1] = ss->ds; out4[2] = ss->dt_s;
    float cfl = sqrtf(ss->T0 / ss->mu) * ss->dt_s / ss->ds;
    out4[3] = cfl * cfl;
}

/* Set-up tables: bridge cells and weights (nb), per-string bridge kernels
 * (ns x nb), neck cells and weights (nn), saddle taps (ns x 4). */
SK_EXPORT void orc_coevo_tables(const OrcCoevo* st, int* bidx, float* bwgt, float* sbk,
                                int* nidx, float* nwgt, int* sidx, float* swgt)
{
    if (st->nb) { memcpy(bidx, st->bidx, st->nb * sizeof(int)); memcpy(bwgt, st->bwgt, st->nb * sizeof(float)); }
    if (st->nb && st->ns && st->sbk) memcpy(sbk, st->sbk, (size_t)st->ns * st->nb * sizeof(float));
    if (st->nn) { memcpy(nidx, st->nidx, st->nn * sizeof(int)); memcpy(nwgt, st->nwgt, st->nn * sizeof(float)); }
    if (st->ns) { memcpy(sidx, st->sad_idx, st->ns * 4 * sizeof(int)); memcpy(swgt, st->sad_wgt, st->ns * 4 * sizeof(float)); }
}
