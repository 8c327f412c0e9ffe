/*
 * omx_oracle.c -- CPU restatement, in plain C and FP64, of the 3D FDTD hot
 * path of One-Mxll-to-Rule-Them-All (src/FDTD).  TEST INFRASTRUCTURE ONLY:
 * it is the checker that the HIP engine is compared against; no product code
 * links, loads or calls it.
 *
 * Arrays keep the reference's Fortran shapes and index bounds (first index
 * fastest) and every loop keeps the reference's loop order and expression
 * order, so that results agree with the reference built from its own sources
 * to rounding.  Each function cites the reference lines it restates.
 *
 * Parity status of each part (how it is pinned):
 *   constants, CPML profiles, den arrays       golden vectors of the reference build
 *   Drude-Lorentz library + coefficients       golden vectors of the reference build
 *   staggered media map                         golden vectors of the reference build
 *   H/E updates, CPML psi, media ADE, sources   golden detector time series of the reference build
 *   detector stencils                           golden detector time series of the reference build
 *   E gather / J scatter (material and single   golden vectors of the reference build (prescribed dP/dt,
 *   groups, several groups)                     q-groups restated in its test driver)
 *   two-level emitter RK4 (no reference code)   closed-form Rabi / free-induction decay in tests
 *   TF/SF plane wave (no reference code)        vacuum leakage and transmitted-amplitude checks in tests
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* constants_mod.f90:15-47 (default-REAL literals kept at their single values) */
static const double pi0 = 3.141592653589793;
static const double c0 = 137.0;
#define EPS0 (1.0 / (4.0 * pi0))
#define MU0 (1.0 / (EPS0 * (c0 * c0)))

typedef struct {
    double *p;
    int l0, l1, l2;      /* lower bounds */
    long e0, e1;         /* extents of the two fastest dims */
    long n;
} arr3;

static void arr_init(arr3* a, int l0, int h0, int l1, int h1, int l2, int h2) {
    a->l0 = l0; a->l1 = l1; a->l2 = l2;
    a->e0 = h0 - l0 + 1; a->e1 = h1 - l1 + 1;
    a->n = a->e0 * a->e1 * (long)(h2 - l2 + 1);
    a->p = (double*)calloc((size_t)a->n, sizeof(double));
}
#define AT(a, i, j, k) ((a).p[((i) - (a).l0) + (a).e0 * (((j) - (a).l1) + (a).e1 * (long)((k) - (a).l2))])

typedef struct {
    int kind, n_poles;
    double eps_r, omega_d, gamma_d;
    double omega_k[8], gamma_k[8], a_k[8];
} orc_medium;   /* same layout as omx_medium */

typedef struct {
    int type;   /* 1 dielectric 2 drude 3 DL */
    int n_poles;
    double A1, A2, C1, C2, C3, C4, eps_r;
    double alpha_k[8], beta_k[8], gamma_k[8], B1_k[8], B2_k[8];
} medium_co;

typedef struct {
    int dir, n_ker, npts;
    double j_amp, w0, t0, tau, t_init, t_final, phase;
    int *ind;        /* npts*3 */
    double *ker;     /* (2n+1)^3 */
    double *jmat;
} source;

typedef struct {
    int type, field, nd;
    int *idx;        /* nd*3 Fortran */
} detector;

typedef struct orc_sim {
    int nx, ny, nz, npml, bc[3];
    double dt, dr, dt_eps0, dt_mu0;
    int cpml_pos[6];
    arr3 Ex, Ey, Ez, Hx, Hy, Hz, Ex_old, Ey_old, Ez_old, eps_x, eps_y, eps_z;
    arr3 Jx, Jy, Jz, Jx_old, Jy_old, Jz_old, dJx, dJy, dJz;
    int have_j;
    double *den_ex, *den_ey, *den_ez, *den_hx, *den_hy, *den_hz;   /* 1-based */
    double be[80], ce[80], bh[80], ch[80];                         /* 1-based */
    arr3 psix_Eyx, psix_Ezx, psiy_Exy, psiy_Ezy, psiz_Exz, psiz_Eyz;
    arr3 psix_Hyx, psix_Hzx, psiy_Hxy, psiy_Hzy, psiz_Hxz, psiz_Hyz;
    /* media */
    int n_media, max_poles;
    medium_co media[16];
    int *media_map;          /* (nx,ny,nz,3) */
    arr3 PDx, PDy, PDz;
    double *PLx, *PLy, *PLz, *PLx_old, *PLy_old, *PLz_old;   /* (nx,ny,nz,P) */
    /* sources, detectors */
    int n_src; source src[16];
    int n_det; detector det[512];
    long n_samples;
    /* quantum sites */
    int nq; int *q_ijk; double q_density, q_dt;
    int q_ngroups; int *q_gtype, *q_gof; double q_w[343];   /* q-groups; single-group kernel weights */
    int n_skip, tq_step_old; double t_skip;   /* quantum sub-stepping (mxim_mxll.f90:73-104) */
    double *tls;             /* 16 per site, same packing as the engine */
    /* tf/sf */
    int tfsf, tf_lo[3], tf_hi[3], tf_kb, tf_n, tf_ks;
    double tf_e0, tf_w0, tf_t0, tf_tau;
    double *ex1, *hy1;
} orc_sim;

/* ------------------------------------------------------------------------ */
/* init_3Dgrid, CPML part (mxll_3D_mod.f90:282-420).  The ratios
 * (npml-kk)/(npml-1.0) are default-REAL and cubed in single precision. */
static void grade_e(int npml_, double sigmaCPML, double sig_to_au, int kk, double* sig, double* alpha, double* kappa) {
    const int npml = npml_;
    float r = (float)(npml - kk) / ((float)npml - 1.0f);
    float r3 = r * r * r;
    float ra = (float)(kk - 1) / ((float)npml - 1.0f);
    *sig = sigmaCPML * (double)r3;
    *alpha = (double)0.05f * sig_to_au * (double)ra;
    *kappa = 1.0 + (5.0 - 1.0) * (double)r3;
}
static void grade_h(int npml_, double sigmaCPML, double sig_to_au, int kk, double* sig, double* alpha, double* kappa) {
    const int npml = npml_;
    float r = ((float)(npml - kk) - 0.5f) / ((float)npml - 1.0f);
    float r3 = r * r * r;
    float ra = ((float)kk - 0.5f) / ((float)npml - 1.0f);
    *sig = sigmaCPML * (double)r3;
    *alpha = (double)0.05f * sig_to_au * (double)ra;
    *kappa = 1.0 + (5.0 - 1.0) * (double)r3;
}

orc_sim* orc_create(int nx, int ny, int nz, int npml, const int* bc, double dt, double dr) {
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->nx = nx; s->ny = ny; s->nz = nz; s->npml = npml;
    s->bc[0] = bc[0]; s->bc[1] = bc[1]; s->bc[2] = bc[2];
    s->dt = dt; s->dr = dr;
    s->dt_eps0 = dt / EPS0;
    s->dt_mu0 = dt / MU0;
    s->n_skip = 1;
    arr_init(&s->Ex, 0, nx, 1, ny + 1, 1, nz + 1);
    arr_init(&s->Ey, 1, nx + 1, 0, ny, 1, nz + 1);
    arr_init(&s->Ez, 1, nx + 1, 1, ny + 1, 0, nz);
    arr_init(&s->Ex_old, 0, nx, 1, ny + 1, 1, nz + 1);
    arr_init(&s->Ey_old, 1, nx + 1, 0, ny, 1, nz + 1);
    arr_init(&s->Ez_old, 1, nx + 1, 1, ny + 1, 0, nz);
    arr_init(&s->Hx, 1, nx, 0, ny, 0, nz);
    arr_init(&s->Hy, 0, nx, 1, ny, 0, nz);
    arr_init(&s->Hz, 0, nx, 0, ny, 1, nz);
    arr_init(&s->eps_x, 1, nx, 1, ny, 1, nz);
    arr_init(&s->eps_y, 1, nx, 1, ny, 1, nz);
    arr_init(&s->eps_z, 1, nx, 1, ny, 1, nz);
    for (long q = 0; q < s->eps_x.n; ++q) s->eps_x.p[q] = s->eps_y.p[q] = s->eps_z.p[q] = 1.0;
    s->den_ex = (double*)calloc(nx + 2, sizeof(double)); s->den_hx = (double*)calloc(nx + 2, sizeof(double));
    s->den_ey = (double*)calloc(ny + 2, sizeof(double)); s->den_hy = (double*)calloc(ny + 2, sizeof(double));
    s->den_ez = (double*)calloc(nz + 2, sizeof(double)); s->den_hz = (double*)calloc(nz + 2, sizeof(double));
    for (int k = 1; k <= nx; ++k) s->den_ex[k] = s->den_hx[k] = 1.0 / dr;
    for (int k = 1; k <= ny; ++k) s->den_ey[k] = s->den_hy[k] = 1.0 / dr;
    for (int k = 1; k <= nz; ++k) s->den_ez[k] = s->den_hz[k] = 1.0 / dr;
    s->media_map = (int*)calloc((size_t)nx * ny * nz * 3, sizeof(int));
    arr_init(&s->PDx, 1, nx, 1, ny, 1, nz);
    arr_init(&s->PDy, 1, nx, 1, ny, 1, nz);
    arr_init(&s->PDz, 1, nx, 1, ny, 1, nz);
    for (int a = 0; a < 3; ++a) s->cpml_pos[2 * a] = s->cpml_pos[2 * a + 1] = (bc[a] == 3);
    const int n_sec = 2;
    if (s->cpml_pos[0]) {
        arr_init(&s->psix_Eyx, 1, n_sec * npml, 1, ny, 1, nz); arr_init(&s->psix_Ezx, 1, n_sec * npml, 1, ny, 1, nz);
        arr_init(&s->psix_Hyx, 1, n_sec * (npml - 1), 1, ny, 1, nz); arr_init(&s->psix_Hzx, 1, n_sec * (npml - 1), 1, ny, 1, nz);
    }
    if (s->cpml_pos[2]) {
        arr_init(&s->psiy_Exy, 1, nx, 1, n_sec * npml, 1, nz); arr_init(&s->psiy_Ezy, 1, nx, 1, n_sec * npml, 1, nz);
        arr_init(&s->psiy_Hxy, 1, nx, 1, n_sec * (npml - 1), 1, nz); arr_init(&s->psiy_Hzy, 1, nx, 1, n_sec * (npml - 1), 1, nz);
    }
    if (s->cpml_pos[4]) {
        arr_init(&s->psiz_Exz, 1, nx, 1, ny, 1, n_sec * npml); arr_init(&s->psiz_Eyz, 1, nx, 1, ny, 1, n_sec * npml);
        arr_init(&s->psiz_Hxz, 1, nx, 1, ny, 1, n_sec * (npml - 1)); arr_init(&s->psiz_Hyz, 1, nx, 1, ny, 1, n_sec * (npml - 1));
    }
    int any = 0;
    for (int f = 0; f < 6; ++f) any |= s->cpml_pos[f];
    if (any) {
        const double sec_to_au = 1.0 / 2.4188843265864e-17, C_to_au = 1.0 / 1.602176634e-19;
        const double kg_to_au = 1.0 / 9.1093837139e-31, m_to_au = 1.0 / 5.29177210544e-11;
        const double sig_to_au = C_to_au * C_to_au * sec_to_au / (kg_to_au * (m_to_au * m_to_au * m_to_au));
        /* 0.8*(m+1) is default REAL */
        const double sigmaCPML = (double)(0.8f * (float)(3 + 1)) / (dr * sqrt(MU0 / EPS0));
        double sig, al, kap;
        for (int kk = 1; kk <= npml; ++kk) {
            grade_e(s->npml, sigmaCPML, sig_to_au, kk, &sig, &al, &kap);
            s->be[kk] = exp(-(sig / kap + al) * dt / EPS0);
            if (sig == 0.0 && al == 0.0 && kk == npml) s->ce[kk] = 0.0;
            else s->ce[kk] = sig * (s->be[kk] - 1.0) / (sig + kap * al) / kap;
        }
        for (int kk = 1; kk <= npml - 1; ++kk) {
            grade_h(s->npml, sigmaCPML, sig_to_au, kk, &sig, &al, &kap);
            s->bh[kk] = exp(-(sig / kap + al) * dt / EPS0);
            s->ch[kk] = sig * (s->bh[kk] - 1.0) / (sig + kap * al) / kap;
        }
        int n[3] = {nx, ny, nz};
        double* de[3] = {s->den_ex, s->den_ey, s->den_ez};
        double* dh[3] = {s->den_hx, s->den_hy, s->den_hz};
        for (int a = 0; a < 3; ++a) {
            if (s->cpml_pos[2 * a]) {
                for (int k = 1; k <= n[a]; ++k)
                    if (k <= npml) { grade_e(s->npml, sigmaCPML, sig_to_au, k, &sig, &al, &kap); de[a][k] = 1.0 / (kap * dr); }
                for (int k = 1; k <= n[a]; ++k)
                    if (k <= npml - 1) { grade_h(s->npml, sigmaCPML, sig_to_au, k, &sig, &al, &kap); dh[a][k] = 1.0 / (kap * dr); }
            }
            if (s->cpml_pos[2 * a + 1]) {
                int kk = npml;
                for (int k = 1; k <= n[a] - 1; ++k)
                    if (k >= n[a] + 1 - npml) { grade_e(s->npml, sigmaCPML, sig_to_au, kk, &sig, &al, &kap); de[a][k] = 1.0 / (kap * dr); kk--; }
                kk = npml - 1;
                for (int k = 1; k <= n[a] - 1; ++k)
                    if (k >= n[a] + 1 - npml) { grade_h(s->npml, sigmaCPML, sig_to_au, kk, &sig, &al, &kap); dh[a][k] = 1.0 / (kap * dr); kk--; }
            }
        }
    }
    return s;
}

void orc_destroy(orc_sim* s) {
    if (!s) return;
    arr3* arrs[] = {&s->Ex, &s->Ey, &s->Ez, &s->Hx, &s->Hy, &s->Hz, &s->Ex_old, &s->Ey_old, &s->Ez_old,
                    &s->eps_x, &s->eps_y, &s->eps_z, &s->Jx, &s->Jy, &s->Jz, &s->Jx_old, &s->Jy_old, &s->Jz_old,
                    &s->dJx, &s->dJy, &s->dJz, &s->psix_Eyx, &s->psix_Ezx, &s->psiy_Exy, &s->psiy_Ezy, &s->psiz_Exz,
                    &s->psiz_Eyz, &s->psix_Hyx, &s->psix_Hzx, &s->psiy_Hxy, &s->psiy_Hzy, &s->psiz_Hxz, &s->psiz_Hyz,
                    &s->PDx, &s->PDy, &s->PDz};
    for (size_t q = 0; q < sizeof(arrs) / sizeof(arrs[0]); ++q) free(arrs[q]->p);
    free(s->den_ex); free(s->den_ey); free(s->den_ez); free(s->den_hx); free(s->den_hy); free(s->den_hz);
    free(s->media_map);
    free(s->PLx); free(s->PLy); free(s->PLz); free(s->PLx_old); free(s->PLy_old); free(s->PLz_old);
    for (int q = 0; q < s->n_src; ++q) { free(s->src[q].ind); free(s->src[q].ker); free(s->src[q].jmat); }
    for (int q = 0; q < s->n_det; ++q) free(s->det[q].idx);
    free(s->q_ijk); free(s->q_gtype); free(s->q_gof); free(s->tls); free(s->ex1); free(s->hy1);
    free(s);
}

void orc_get_cpml(const orc_sim* s, double* be, double* ce, double* bh, double* ch, double* den) {
    for (int k = 1; k <= s->npml; ++k) { be[k - 1] = s->be[k]; ce[k - 1] = s->ce[k]; }
    for (int k = 1; k <= s->npml - 1; ++k) { bh[k - 1] = s->bh[k]; ch[k - 1] = s->ch[k]; }
    long w = 0;
    const double* arrs[6] = {s->den_ex, s->den_ey, s->den_ez, s->den_hx, s->den_hy, s->den_hz};
    const int n[3] = {s->nx, s->ny, s->nz};
    for (int q = 0; q < 6; ++q) for (int k = 1; k <= n[q % 3]; ++k) den[w++] = arrs[q][k];
}

/* ------------------------------------------------------------------------ */
/* get_DL_library_data (classical_medium_mod.f90:125-199); literals are REAL(4). */
int orc_library(const char* name, double eps_r, orc_medium* m) {
    const double ev = (double)(1.0f / 27.2114f);
    memset(m, 0, sizeof(*m));
    m->kind = 3; m->eps_r = eps_r;
    double wp;
    if (!strcmp(name, "Ag")) {
        const float w[5] = {0.816f, 4.481f, 8.185f, 9.083f, 20.29f}, g[5] = {3.886f, 0.452f, 0.065f, 0.916f, 2.419f};
        const float f[5] = {0.065f, 0.124f, 0.011f, 0.840f, 5.646f};
        m->n_poles = 5; m->gamma_d = (double)0.048f * ev; wp = (double)9.01f * ev;
        m->omega_d = (double)(9.01f * 0.919238815542512f) * ev;
        for (int q = 0; q < 5; ++q) { m->omega_k[q] = (double)w[q] * ev; m->gamma_k[q] = (double)g[q] * ev; m->a_k[q] = (double)f[q] * (wp * wp); }
    } else if (!strcmp(name, "Al")) {
        const float w[4] = {0.162f, 1.544f, 1.808f, 3.473f}, g[4] = {0.333f, 0.312f, 1.351f, 3.382f};
        const float f[4] = {0.227f, 0.050f, 0.166f, 0.030f};
        m->n_poles = 4; m->gamma_d = (double)0.047f * ev; wp = (double)14.98f * ev;
        m->omega_d = (double)(14.98f * 0.723187389270582f) * ev;
        for (int q = 0; q < 4; ++q) { m->omega_k[q] = (double)w[q] * ev; m->gamma_k[q] = (double)g[q] * ev; m->a_k[q] = (double)f[q] * (wp * wp); }
    } else if (!strcmp(name, "Au")) {
        const float w[5] = {0.415f, 0.830f, 2.969f, 4.304f, 13.32f}, g[5] = {0.241f, 0.345f, 0.870f, 2.494f, 2.214f};
        const float f[5] = {0.024f, 0.010f, 0.071f, 0.601f, 4.384f};
        m->n_poles = 5; m->gamma_d = (double)0.053f * ev; wp = (double)9.03f * ev;
        m->omega_d = (double)(9.03f * 0.871779788708135f) * ev;
        for (int q = 0; q < 5; ++q) { m->omega_k[q] = (double)w[q] * ev; m->gamma_k[q] = (double)g[q] * ev; m->a_k[q] = (double)f[q] * (wp * wp); }
    } else {
        return 1;
    }
    return 0;
}

/* init_classical_medium (classical_medium_mod.f90:46-101). */
static void medium_coeffs(const orc_medium* in, double dt, medium_co* c) {
    memset(c, 0, sizeof(*c));
    c->type = in->kind; c->eps_r = in->eps_r; c->n_poles = in->n_poles;
    if (in->kind == 2) {
        c->A1 = (2.0 - in->gamma_d * dt) / (2.0 + in->gamma_d * dt);
        c->A2 = EPS0 * (in->omega_d * in->omega_d) * dt / (2.0 + in->gamma_d * dt);
        c->C1 = (c->eps_r * EPS0 / dt - 0.5 * c->A2) / (c->eps_r * EPS0 / dt + 0.5 * c->A2);
        c->C3 = 1.0 / (c->eps_r * EPS0 / dt + 0.5 * c->A2);
        c->C4 = 0.5 * (c->A1 + 1.0) / (c->eps_r * EPS0 / dt + 0.5 * c->A2);
    } else if (in->kind == 3) {
        for (int i = 0; i < in->n_poles; ++i) {
            c->alpha_k[i] = (2.0 - (in->omega_k[i] * in->omega_k[i]) * (dt * dt)) / (1.0 + 0.5 * in->gamma_k[i] * dt);
            c->beta_k[i] = (in->gamma_k[i] * dt - 2.0) / (in->gamma_k[i] * dt + 2.0);
            c->gamma_k[i] = EPS0 * in->a_k[i] * dt / (in->gamma_k[i] * dt + 2.0);
        }
        double S = 0.0;
        for (int i = 0; i < in->n_poles; ++i) S += c->gamma_k[i];
        c->A1 = (2.0 - in->gamma_d * dt) / (2.0 + in->gamma_d * dt);
        c->A2 = EPS0 * (in->omega_d * in->omega_d) * dt / (2.0 + in->gamma_d * dt);
        const double D = c->eps_r * EPS0 / dt + 0.5 * c->A2 + 0.5 * S;
        c->C1 = (c->eps_r * EPS0 / dt - 0.5 * c->A2) / D;
        c->C2 = 0.5 * S / D;
        c->C3 = 1.0 / D;
        c->C4 = 0.5 * (c->A1 + 1.0) / D;
        for (int i = 0; i < in->n_poles; ++i) {
            c->B1_k[i] = 0.5 * (1.0 + c->alpha_k[i]) / D;
            c->B2_k[i] = 0.5 * c->beta_k[i] / D;
        }
    }
}

/* read_init_media_3D after the point lists are read (classical_medium_mod.f90:694-765):
 * aux is the cell medium map (nx,ny,nz), Fortran order.  Neighbours past the
 * grid read as 0 (the reference reads past its array there). */
int orc_set_media(orc_sim* s, int n_media, const orc_medium* table, const int8_t* aux) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    s->n_media = n_media;
    s->max_poles = 1;
    for (int n = 1; n <= n_media; ++n) {
        medium_coeffs(&table[n - 1], s->dt, &s->media[n]);
        if (table[n - 1].kind == 3 && table[n - 1].n_poles > s->max_poles) s->max_poles = table[n - 1].n_poles;
    }
#define AUX(i, j, k) (((i) > nx || (j) > ny || (k) > nz) ? 0 : (int)aux[((i) - 1) + (long)nx * (((j) - 1) + (long)ny * ((k) - 1))])
#define MAP(i, j, k, c) s->media_map[((i) - 1) + (long)nx * (((j) - 1) + (long)ny * (((k) - 1) + (long)nz * (c)))]
    for (int n = 1; n <= n_media; ++n)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    if (i == nx && AUX(i, j, k) == n) MAP(i, j, k, 0) = n;
                    else if (j == ny && AUX(i, j, k) == n) MAP(i, j, k, 0) = n;
                    else if (k == nz && AUX(i, j, k) == n) MAP(i, j, k, 0) = n;
                    else {
                        if (AUX(i, j, k) == n && AUX(i + 1, j, k) == n) MAP(i, j, k, 0) = n;
                        else if (AUX(i, j, k) != 0 && AUX(i + 1, j, k) == n) MAP(i, j, k, 0) = n;
                        if (AUX(i, j, k) == n && AUX(i, j + 1, k) == n) MAP(i, j, k, 1) = n;
                        else if (AUX(i, j, k) != 0 && AUX(i, j + 1, k) == n) MAP(i, j, k, 1) = n;
                        if (AUX(i, j, k) == n && AUX(i, j, k + 1) == n) MAP(i, j, k, 2) = n;
                        else if (AUX(i, j, k) != 0 && AUX(i, j, k + 1) == n) MAP(i, j, k, 2) = n;
                    }
                }
    /* dielectric: eps = 1/eps_r (the reference leaves eps_r unset here; the
     * engine and this restatement both use the file's value) */
    for (int n = 1; n <= n_media; ++n) {
        if (s->media[n].type != 1) continue;
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    if (MAP(i, j, k, 0) == n) AT(s->eps_x, i, j, k) = 1.0 / s->media[n].eps_r;
                    if (MAP(i, j, k, 1) == n) AT(s->eps_y, i, j, k) = 1.0 / s->media[n].eps_r;
                    if (MAP(i, j, k, 2) == n) AT(s->eps_z, i, j, k) = 1.0 / s->media[n].eps_r;
                }
    }
#undef AUX
    const size_t np = (size_t)nx * ny * nz * s->max_poles;
    s->PLx = (double*)calloc(np, sizeof(double)); s->PLy = (double*)calloc(np, sizeof(double));
    s->PLz = (double*)calloc(np, sizeof(double)); s->PLx_old = (double*)calloc(np, sizeof(double));
    s->PLy_old = (double*)calloc(np, sizeof(double)); s->PLz_old = (double*)calloc(np, sizeof(double));
    return 0;
}

void orc_get_media_map(const orc_sim* s, int32_t* out) {
    memcpy(out, s->media_map, sizeof(int) * (size_t)s->nx * s->ny * s->nz * 3);
}

void orc_get_medium_coeffs(const orc_sim* s, int m, double* out) {
    const medium_co* c = &s->media[m];
    out[0] = c->A1; out[1] = c->A2; out[2] = c->C1; out[3] = c->C2; out[4] = c->C3; out[5] = c->C4;
    out[6] = c->type == 1 ? 1.0 / c->eps_r : 1.0;
    for (int q = 0; q < 8; ++q) {
        out[7 + q] = c->alpha_k[q]; out[15 + q] = c->beta_k[q]; out[23 + q] = c->gamma_k[q];
        out[31 + q] = c->B1_k[q]; out[39 + q] = c->B2_k[q];
    }
}

/* get_medium_polarization (classical_medium_mod.f90:773-833) */
static int medium_polarization(const medium_co* media, int idx, double* PDi, double* PLi, double* PLi_old, long pstride,
                               double* Ei, double* Ei_old, double rotH) {
    if (idx == 0) return 1;
    const medium_co* m = &media[idx];
    if (m->type == 2) {
        double tmpE = m->C1 * *Ei + m->C3 * rotH - m->C4 * *PDi;
        *PDi = m->A1 * *PDi + m->A2 * (tmpE + *Ei);
        *Ei = tmpE;
        return 0;
    }
    if (m->type == 3) {
        double tmpPL[8];
        double tmpE = m->C1 * *Ei + m->C2 * *Ei_old + m->C3 * rotH - m->C4 * *PDi;
        for (int i = 0; i < m->n_poles; ++i) tmpE = tmpE - m->B1_k[i] * PLi[i * pstride] - m->B2_k[i] * PLi_old[i * pstride];
        *PDi = m->A1 * *PDi + m->A2 * (tmpE + *Ei);
        for (int i = 0; i < m->n_poles; ++i)
            tmpPL[i] = m->alpha_k[i] * PLi[i * pstride] + m->beta_k[i] * PLi_old[i * pstride] + m->gamma_k[i] * (tmpE - *Ei_old);
        *Ei_old = *Ei;
        for (int i = 0; i < m->n_poles; ++i) { PLi_old[i * pstride] = PLi[i * pstride]; PLi[i * pstride] = tmpPL[i]; }
        *Ei = tmpE;
        return 0;
    }
    return 1;   /* dielectric */
}

/* ------------------------------------------------------------------------ */
/* td_propagate_H_3Dfield (mxll_3D_mod.f90:505-696) */
void orc_step_h(orc_sim* s) {
    const int nx = s->nx, ny = s->ny, nz = s->nz, npml = s->npml;
    const int n_base = npml - 1;
    const double dt_mu0 = s->dt_mu0, dr = s->dr;
    if (s->bc[0] == 2) {
        for (int k = 0; k <= nz; ++k) for (int j = 1; j <= ny + 1; ++j) AT(s->Ez, nx + 1, j, k) = AT(s->Ez, 1, j, k);
        for (int k = 1; k <= nz + 1; ++k) for (int j = 0; j <= ny; ++j) AT(s->Ey, nx + 1, j, k) = AT(s->Ey, 1, j, k);
    }
    if (s->bc[1] == 2) {
        for (int k = 1; k <= nz + 1; ++k) for (int i = 0; i <= nx; ++i) AT(s->Ex, i, ny + 1, k) = AT(s->Ex, i, 1, k);
        for (int k = 0; k <= nz; ++k) for (int i = 1; i <= nx + 1; ++i) AT(s->Ez, i, ny + 1, k) = AT(s->Ez, i, 1, k);
    }
    if (s->bc[2] == 2) {
        for (int j = 1; j <= ny + 1; ++j) for (int i = 0; i <= nx; ++i) AT(s->Ex, i, j, nz + 1) = AT(s->Ex, i, j, 1);
        for (int j = 0; j <= ny; ++j) for (int i = 1; i <= nx + 1; ++i) AT(s->Ey, i, j, nz + 1) = AT(s->Ey, i, j, 1);
    }
    /* the reference runs i, j, k with k innermost (mxll_3D_mod.f90:548-550);
     * every cell is independent, so the order here (i innermost, the arrays'
     * fastest index) gives the same values */
#pragma omp parallel
    {
#pragma omp for collapse(2) schedule(static) nowait
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double rotE = (AT(s->Ez, i, j, k) - AT(s->Ez, i, j + 1, k)) * s->den_hy[j] +
                                  (AT(s->Ey, i, j, k + 1) - AT(s->Ey, i, j, k)) * s->den_hz[k];
                    AT(s->Hx, i, j, k) = AT(s->Hx, i, j, k) + dt_mu0 * rotE;
                }
#pragma omp for collapse(2) schedule(static) nowait
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double rotE = (AT(s->Ez, i + 1, j, k) - AT(s->Ez, i, j, k)) * s->den_hx[i] +
                                  (AT(s->Ex, i, j, k) - AT(s->Ex, i, j, k + 1)) * s->den_hz[k];
                    AT(s->Hy, i, j, k) = AT(s->Hy, i, j, k) + dt_mu0 * rotE;
                }
#pragma omp for collapse(2) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double rotE = (AT(s->Ey, i, j, k) - AT(s->Ey, i + 1, j, k)) * s->den_hx[i] +
                                  (AT(s->Ex, i, j + 1, k) - AT(s->Ex, i, j, k)) * s->den_hy[j];
                    AT(s->Hz, i, j, k) = AT(s->Hz, i, j, k) + dt_mu0 * rotE;
                }
    }
    if (s->cpml_pos[0]) {
        #pragma omp parallel for schedule(static)
        for (int j = 1; j <= ny; ++j) for (int k = 1; k <= nz; ++k) for (int i = 1; i <= npml - 1; ++i) {
            AT(s->psix_Hyx, i, j, k) = s->bh[i] * AT(s->psix_Hyx, i, j, k) + s->ch[i] * (AT(s->Ez, i + 1, j, k) - AT(s->Ez, i, j, k)) / dr;
            AT(s->psix_Hzx, i, j, k) = s->bh[i] * AT(s->psix_Hzx, i, j, k) + s->ch[i] * (AT(s->Ey, i, j, k) - AT(s->Ey, i + 1, j, k)) / dr;
            AT(s->Hy, i, j, k) = AT(s->Hy, i, j, k) + dt_mu0 * AT(s->psix_Hyx, i, j, k);
            AT(s->Hz, i, j, k) = AT(s->Hz, i, j, k) + dt_mu0 * AT(s->psix_Hzx, i, j, k);
        }
    }
    if (s->cpml_pos[1]) {
        #pragma omp parallel for schedule(static)
        for (int j = 1; j <= ny; ++j) for (int k = 1; k <= nz; ++k) {
            int ii = 2 * (npml - 1);
            for (int i = nx + 1 - npml; i <= nx - 1; ++i) {
                AT(s->psix_Hyx, ii, j, k) = s->bh[ii - n_base] * AT(s->psix_Hyx, ii, j, k) + s->ch[ii - n_base] * (AT(s->Ez, i + 1, j, k) - AT(s->Ez, i, j, k)) / dr;
                AT(s->psix_Hzx, ii, j, k) = s->bh[ii - n_base] * AT(s->psix_Hzx, ii, j, k) + s->ch[ii - n_base] * (AT(s->Ey, i, j, k) - AT(s->Ey, i + 1, j, k)) / dr;
                AT(s->Hy, i, j, k) = AT(s->Hy, i, j, k) + dt_mu0 * AT(s->psix_Hyx, ii, j, k);
                AT(s->Hz, i, j, k) = AT(s->Hz, i, j, k) + dt_mu0 * AT(s->psix_Hzx, ii, j, k);
                ii = ii - 1;
            }
        }
    }
    if (s->cpml_pos[2]) {
        #pragma omp parallel for schedule(static)
        for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) for (int j = 1; j <= npml - 1; ++j) {
            AT(s->psiy_Hxy, i, j, k) = s->bh[j] * AT(s->psiy_Hxy, i, j, k) + s->ch[j] * (AT(s->Ez, i, j, k) - AT(s->Ez, i, j + 1, k)) / dr;
            AT(s->psiy_Hzy, i, j, k) = s->bh[j] * AT(s->psiy_Hzy, i, j, k) + s->ch[j] * (AT(s->Ex, i, j + 1, k) - AT(s->Ex, i, j, k)) / dr;
            AT(s->Hx, i, j, k) = AT(s->Hx, i, j, k) + dt_mu0 * AT(s->psiy_Hxy, i, j, k);
            AT(s->Hz, i, j, k) = AT(s->Hz, i, j, k) + dt_mu0 * AT(s->psiy_Hzy, i, j, k);
        }
    }
    if (s->cpml_pos[3]) {
        #pragma omp parallel for schedule(static)
        for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) {
            int jj = 2 * (npml - 1);
            for (int j = ny + 1 - npml; j <= ny - 1; ++j) {
                AT(s->psiy_Hxy, i, jj, k) = s->bh[jj - n_base] * AT(s->psiy_Hxy, i, jj, k) + s->ch[jj - n_base] * (AT(s->Ez, i, j, k) - AT(s->Ez, i, j + 1, k)) / dr;
                AT(s->psiy_Hzy, i, jj, k) = s->bh[jj - n_base] * AT(s->psiy_Hzy, i, jj, k) + s->ch[jj - n_base] * (AT(s->Ex, i, j + 1, k) - AT(s->Ex, i, j, k)) / dr;
                AT(s->Hx, i, j, k) = AT(s->Hx, i, j, k) + dt_mu0 * AT(s->psiy_Hxy, i, jj, k);
                AT(s->Hz, i, j, k) = AT(s->Hz, i, j, k) + dt_mu0 * AT(s->psiy_Hzy, i, jj, k);
                jj = jj - 1;
            }
        }
    }
    if (s->cpml_pos[4]) {
        #pragma omp parallel for schedule(static)
        for (int i = 1; i <= nx; ++i) for (int j = 1; j <= ny; ++j) for (int k = 1; k <= npml - 1; ++k) {
            AT(s->psiz_Hxz, i, j, k) = s->bh[k] * AT(s->psiz_Hxz, i, j, k) + s->ch[k] * (AT(s->Ey, i, j, k + 1) - AT(s->Ey, i, j, k)) / dr;
            AT(s->psiz_Hyz, i, j, k) = s->bh[k] * AT(s->psiz_Hyz, i, j, k) + s->ch[k] * (AT(s->Ex, i, j, k) - AT(s->Ex, i, j, k + 1)) / dr;
            AT(s->Hx, i, j, k) = AT(s->Hx, i, j, k) + dt_mu0 * AT(s->psiz_Hxz, i, j, k);
            AT(s->Hy, i, j, k) = AT(s->Hy, i, j, k) + dt_mu0 * AT(s->psiz_Hyz, i, j, k);
        }
    }
    if (s->cpml_pos[5]) {
        #pragma omp parallel for schedule(static)
        for (int i = 1; i <= nx; ++i) for (int j = 1; j <= ny; ++j) {
            int kk = 2 * (npml - 1);
            for (int k = nz + 1 - npml; k <= nz - 1; ++k) {
                AT(s->psiz_Hxz, i, j, kk) = s->bh[kk - n_base] * AT(s->psiz_Hxz, i, j, kk) + s->ch[kk - n_base] * (AT(s->Ey, i, j, k + 1) - AT(s->Ey, i, j, k)) / dr;
                AT(s->psiz_Hyz, i, j, kk) = s->bh[kk - n_base] * AT(s->psiz_Hyz, i, j, kk) + s->ch[kk - n_base] * (AT(s->Ex, i, j, k) - AT(s->Ex, i, j, k + 1)) / dr;
                AT(s->Hx, i, j, k) = AT(s->Hx, i, j, k) + dt_mu0 * AT(s->psiz_Hxz, i, j, kk);
                AT(s->Hy, i, j, k) = AT(s->Hy, i, j, k) + dt_mu0 * AT(s->psiz_Hyz, i, j, kk);
                kk = kk - 1;
            }
        }
    }
}

/* td_propagate_E_3Dfield (mxll_3D_mod.f90:700-923) */
void orc_step_e(orc_sim* s, double time, double t_skip) {
    const int nx = s->nx, ny = s->ny, nz = s->nz, npml = s->npml;
    const int n_base = npml;
    const double dt_eps0 = s->dt_eps0, dr = s->dr, tau = time - t_skip;
    const long pst = (long)nx * ny * nz;
    if (s->bc[0] == 2) {
        for (int k = 0; k <= nz; ++k) for (int j = 1; j <= ny; ++j) AT(s->Hy, 0, j, k) = AT(s->Hy, nx, j, k);
        for (int k = 1; k <= nz; ++k) for (int j = 0; j <= ny; ++j) AT(s->Hz, 0, j, k) = AT(s->Hz, nx, j, k);
    }
    if (s->bc[1] == 2) {
        for (int k = 0; k <= nz; ++k) for (int i = 1; i <= nx; ++i) AT(s->Hx, i, 0, k) = AT(s->Hx, i, ny, k);
        for (int k = 1; k <= nz; ++k) for (int i = 0; i <= nx; ++i) AT(s->Hz, i, 0, k) = AT(s->Hz, i, ny, k);
    }
    if (s->bc[2] == 2) {
        for (int j = 0; j <= ny; ++j) for (int i = 1; i <= nx; ++i) AT(s->Hx, i, j, 0) = AT(s->Hx, i, j, nz);
        for (int j = 1; j <= ny; ++j) for (int i = 0; i <= nx; ++i) AT(s->Hy, i, j, 0) = AT(s->Hy, i, j, nz);
    }
    const int hj = s->have_j;
#define MM(i, j, k, c) s->media_map[((i) - 1) + (long)nx * (((j) - 1) + (long)ny * (((k) - 1) + (long)nz * (c)))]
#define PIDX(i, j, k) (((i) - 1) + (long)nx * (((j) - 1) + (long)ny * ((k) - 1)))
#pragma omp parallel
    {
#pragma omp for collapse(3) schedule(static) nowait
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double rotH = (AT(s->Hz, i, j, k) - AT(s->Hz, i, j - 1, k)) * s->den_ey[j] +
                                  (AT(s->Hy, i, j, k - 1) - AT(s->Hy, i, j, k)) * s->den_ez[k];
                    const long p = PIDX(i, j, k);
                    int no_medium = medium_polarization(s->media, MM(i, j, k, 0), &AT(s->PDx, i, j, k), s->PLx ? s->PLx + p : 0,
                                                        s->PLx_old ? s->PLx_old + p : 0, pst, &AT(s->Ex, i, j, k), &AT(s->Ex_old, i, j, k), rotH);
                    if (no_medium) {
                        double Jav = 0.0;
                        if (hj) Jav = 0.5 * (AT(s->Jx_old, i, j, k) + tau * AT(s->dJx, i, j, k) + AT(s->Jx_old, i + 1, j, k) + tau * AT(s->dJx, i + 1, j, k));
                        AT(s->Ex, i, j, k) = AT(s->Ex, i, j, k) + dt_eps0 * AT(s->eps_x, i, j, k) * rotH - dt_eps0 * Jav;
                    }
                }
#pragma omp for collapse(3) schedule(static) nowait
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double rotH = (AT(s->Hz, i - 1, j, k) - AT(s->Hz, i, j, k)) * s->den_ex[i] +
                                  (AT(s->Hx, i, j, k) - AT(s->Hx, i, j, k - 1)) * s->den_ez[k];
                    const long p = PIDX(i, j, k);
                    int no_medium = medium_polarization(s->media, MM(i, j, k, 1), &AT(s->PDy, i, j, k), s->PLy ? s->PLy + p : 0,
                                                        s->PLy_old ? s->PLy_old + p : 0, pst, &AT(s->Ey, i, j, k), &AT(s->Ey_old, i, j, k), rotH);
                    if (no_medium) {
                        double Jav = 0.0;
                        if (hj) Jav = 0.5 * (AT(s->Jy_old, i, j, k) + tau * AT(s->dJy, i, j, k) + AT(s->Jy_old, i, j + 1, k) + tau * AT(s->dJy, i, j + 1, k));
                        AT(s->Ey, i, j, k) = AT(s->Ey, i, j, k) + dt_eps0 * AT(s->eps_y, i, j, k) * rotH - dt_eps0 * Jav;
                    }
                }
#pragma omp for collapse(3) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    double rotH = (AT(s->Hy, i, j, k) - AT(s->Hy, i - 1, j, k)) * s->den_ex[i] +
                                  (AT(s->Hx, i, j - 1, k) - AT(s->Hx, i, j, k)) * s->den_ey[j];
                    const long p = PIDX(i, j, k);
                    int no_medium = medium_polarization(s->media, MM(i, j, k, 2), &AT(s->PDz, i, j, k), s->PLz ? s->PLz + p : 0,
                                                        s->PLz_old ? s->PLz_old + p : 0, pst, &AT(s->Ez, i, j, k), &AT(s->Ez_old, i, j, k), rotH);
                    if (no_medium) {
                        double Jav = 0.0;
                        if (hj) Jav = 0.5 * (AT(s->Jz_old, i, j, k) + tau * AT(s->dJz, i, j, k) + AT(s->Jz_old, i, j, k + 1) + tau * AT(s->dJz, i, j, k + 1));
                        AT(s->Ez, i, j, k) = AT(s->Ez, i, j, k) + dt_eps0 * AT(s->eps_z, i, j, k) * rotH - dt_eps0 * Jav;
                    }
                }
    }
#undef MM
#undef PIDX
    if (s->cpml_pos[0]) {
        #pragma omp parallel for schedule(static)
        for (int j = 1; j <= ny; ++j) for (int k = 1; k <= nz; ++k) for (int i = 2; i <= npml; ++i) {
            AT(s->psix_Eyx, i, j, k) = s->be[i] * AT(s->psix_Eyx, i, j, k) + s->ce[i] * (AT(s->Hz, i - 1, j, k) - AT(s->Hz, i, j, k)) / dr;
            AT(s->psix_Ezx, i, j, k) = s->be[i] * AT(s->psix_Ezx, i, j, k) + s->ce[i] * (AT(s->Hy, i, j, k) - AT(s->Hy, i - 1, j, k)) / dr;
            AT(s->Ey, i, j, k) = AT(s->Ey, i, j, k) + dt_eps0 * AT(s->psix_Eyx, i, j, k);
            AT(s->Ez, i, j, k) = AT(s->Ez, i, j, k) + dt_eps0 * AT(s->psix_Ezx, i, j, k);
        }
    }
    if (s->cpml_pos[1]) {
        #pragma omp parallel for schedule(static)
        for (int j = 1; j <= ny; ++j) for (int k = 1; k <= nz; ++k) {
            int ii = 2 * npml;
            for (int i = nx + 1 - npml; i <= nx - 1; ++i) {
                AT(s->psix_Eyx, ii, j, k) = s->be[ii - n_base] * AT(s->psix_Eyx, ii, j, k) + s->ce[ii - n_base] * (AT(s->Hz, i - 1, j, k) - AT(s->Hz, i, j, k)) / dr;
                AT(s->psix_Ezx, ii, j, k) = s->be[ii - n_base] * AT(s->psix_Ezx, ii, j, k) + s->ce[ii - n_base] * (AT(s->Hy, i, j, k) - AT(s->Hy, i - 1, j, k)) / dr;
                AT(s->Ey, i, j, k) = AT(s->Ey, i, j, k) + dt_eps0 * AT(s->psix_Eyx, ii, j, k);
                AT(s->Ez, i, j, k) = AT(s->Ez, i, j, k) + dt_eps0 * AT(s->psix_Ezx, ii, j, k);
                ii = ii - 1;
            }
        }
    }
    if (s->cpml_pos[2]) {
        #pragma omp parallel for schedule(static)
        for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) for (int j = 2; j <= npml; ++j) {
            AT(s->psiy_Exy, i, j, k) = s->be[j] * AT(s->psiy_Exy, i, j, k) + s->ce[j] * (AT(s->Hz, i, j, k) - AT(s->Hz, i, j - 1, k)) / dr;
            AT(s->psiy_Ezy, i, j, k) = s->be[j] * AT(s->psiy_Ezy, i, j, k) + s->ce[j] * (AT(s->Hx, i, j - 1, k) - AT(s->Hx, i, j, k)) / dr;
            AT(s->Ex, i, j, k) = AT(s->Ex, i, j, k) + dt_eps0 * AT(s->psiy_Exy, i, j, k);
            AT(s->Ez, i, j, k) = AT(s->Ez, i, j, k) + dt_eps0 * AT(s->psiy_Ezy, i, j, k);
        }
    }
    if (s->cpml_pos[3]) {
        #pragma omp parallel for schedule(static)
        for (int k = 1; k <= nz; ++k) for (int i = 1; i <= nx; ++i) {
            int jj = 2 * npml;
            for (int j = ny + 1 - npml; j <= ny - 1; ++j) {
                AT(s->psiy_Exy, i, jj, k) = s->be[jj - n_base] * AT(s->psiy_Exy, i, jj, k) + s->ce[jj - n_base] * (AT(s->Hz, i, j, k) - AT(s->Hz, i, j - 1, k)) / dr;
                AT(s->psiy_Ezy, i, jj, k) = s->be[jj - n_base] * AT(s->psiy_Ezy, i, jj, k) + s->ce[jj - n_base] * (AT(s->Hx, i, j - 1, k) - AT(s->Hx, i, j, k)) / dr;
                AT(s->Ex, i, j, k) = AT(s->Ex, i, j, k) + dt_eps0 * AT(s->psiy_Exy, i, jj, k);
                AT(s->Ez, i, j, k) = AT(s->Ez, i, j, k) + dt_eps0 * AT(s->psiy_Ezy, i, jj, k);
                jj = jj - 1;
            }
        }
    }
    if (s->cpml_pos[4]) {
        #pragma omp parallel for schedule(static)
        for (int i = 1; i <= nx; ++i) for (int j = 1; j <= ny; ++j) for (int k = 2; k <= npml; ++k) {
            AT(s->psiz_Exz, i, j, k) = s->be[k] * AT(s->psiz_Exz, i, j, k) + s->ce[k] * (AT(s->Hy, i, j, k - 1) - AT(s->Hy, i, j, k)) / dr;
            AT(s->psiz_Eyz, i, j, k) = s->be[k] * AT(s->psiz_Eyz, i, j, k) + s->ce[k] * (AT(s->Hx, i, j, k) - AT(s->Hx, i, j, k - 1)) / dr;
            AT(s->Ex, i, j, k) = AT(s->Ex, i, j, k) + dt_eps0 * AT(s->psiz_Exz, i, j, k);
            AT(s->Ey, i, j, k) = AT(s->Ey, i, j, k) + dt_eps0 * AT(s->psiz_Eyz, i, j, k);
        }
    }
    if (s->cpml_pos[5]) {
        #pragma omp parallel for schedule(static)
        for (int i = 1; i <= nx; ++i) for (int j = 1; j <= ny; ++j) {
            int kk = 2 * npml;
            for (int k = nz + 1 - npml; k <= nz - 1; ++k) {
                AT(s->psiz_Exz, i, j, kk) = s->be[kk - n_base] * AT(s->psiz_Exz, i, j, kk) + s->ce[kk - n_base] * (AT(s->Hy, i, j, k - 1) - AT(s->Hy, i, j, k)) / dr;
                AT(s->psiz_Eyz, i, j, kk) = s->be[kk - n_base] * AT(s->psiz_Eyz, i, j, kk) + s->ce[kk - n_base] * (AT(s->Hx, i, j, k) - AT(s->Hx, i, j, k - 1)) / dr;
                AT(s->Ex, i, j, k) = AT(s->Ex, i, j, k) + dt_eps0 * AT(s->psiz_Exz, i, j, kk);
                AT(s->Ey, i, j, k) = AT(s->Ey, i, j, k) + dt_eps0 * AT(s->psiz_Eyz, i, j, kk);
                kk = kk - 1;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* compute_kernel, 3D branch (source_mod.f90:251-327) */
void orc_source_kernel(int n_ker, double dr, double* ker) {
    const double aBH[4] = {0.353222222, -0.488, 0.145, -0.010222222};
    const double r_max = (n_ker * 2 + 2) * dr, r0 = r_max / 2.0;
    const int w = 2 * n_ker + 1;
    double norm = 0.0;
    for (int k = -n_ker; k <= n_ker; ++k)
        for (int j = -n_ker; j <= n_ker; ++j)
            for (int i = -n_ker; i <= n_ker; ++i) {
                const double x = i * dr, y = j * dr, z = k * dr;
                const double r = sqrt(x * x + y * y + z * z);
                double v = aBH[0] + aBH[1] * cos(2.0 * pi0 * (r - r0) / r_max) + aBH[2] * cos(2.0 * pi0 * 2.0 * (r - r0) / r_max) +
                           aBH[3] * cos(2.0 * pi0 * 3.0 * (r - r0) / r_max);
                if (v < 0.0) v = 0.0;
                ker[(i + n_ker) + w * ((j + n_ker) + w * (k + n_ker))] = v;
                norm = norm + v;
            }
    for (long q = 0; q < (long)w * w * w; ++q) ker[q] = ker[q] / norm;
}

/* determine_indx_and_ranks, 3D non-MPI (source_mod.f90:347-356); r0 in au. */
void orc_source_indices(int n_ker, const double* r0, double dr, const int* grid, int* ind) {
    const int w = 2 * n_ker + 1;
    const long np = (long)w * w * w;
    for (int k = -n_ker; k <= n_ker; ++k)
        for (int j = -n_ker; j <= n_ker; ++j)
            for (int i = -n_ker; i <= n_ker; ++i) {
                const long p = (i + n_ker) + w * ((j + n_ker) + w * (long)(k + n_ker));
                ind[p] = i + (int)(r0[0] / dr) + grid[0] / 2;
                ind[p + np] = j + (int)(r0[1] / dr) + grid[1] / 2;
                ind[p + 2 * np] = k + (int)(r0[2] / dr) + grid[2] / 2;
            }
}

int orc_add_source(orc_sim* s, int dir, double j_amp, double w0, double t0, double tau, double t_init, double t_final,
                   double phase, int n_ker, const int* ind, const double* ker) {
    source* q = &s->src[s->n_src++];
    const int w = 2 * n_ker + 1;
    const long np = (long)w * w * w;
    q->dir = dir; q->n_ker = n_ker; q->npts = (int)np;
    q->j_amp = j_amp; q->w0 = w0; q->t0 = t0; q->tau = tau; q->t_init = t_init; q->t_final = t_final; q->phase = phase;
    q->ind = (int*)malloc(sizeof(int) * 3 * np); memcpy(q->ind, ind, sizeof(int) * 3 * np);
    q->ker = (double*)malloc(sizeof(double) * np); memcpy(q->ker, ker, sizeof(double) * np);
    q->jmat = (double*)calloc((size_t)np, sizeof(double));
    return 0;
}

/* update_source (source_mod.f90:106-132) + source_interactions 3D
 * (interactions_mod.f90:102-156); J_mat one step past the kernel reads 0. */
void orc_inject_sources(orc_sim* s, double time) {
    const double c_src = s->dt_eps0 / s->dr / c0 / 2.0;
    for (int q = 0; q < s->n_src; ++q) {
        source* S = &s->src[q];
        const int n = S->n_ker, w = 2 * n + 1;
        const long np = (long)w * w * w;
        if (time >= S->t_init && time <= S->t_final) {
            const double x = (time - S->t0) / S->tau;
            const double envelope = exp(-(x * x));
            const double cos_t = cos(S->w0 * (time - S->t0) + S->phase);
            for (long p = 0; p < np; ++p) S->jmat[p] = S->j_amp * envelope * cos_t * S->ker[p];
        } else {
            for (long p = 0; p < np; ++p) S->jmat[p] = 0.0;
        }
        for (int i = -n; i <= n; ++i)
            for (int j = -n; j <= n; ++j)
                for (int k = -n; k <= n; ++k) {
                    const long p = (i + n) + w * ((j + n) + w * (long)(k + n));
                    const int ii = S->ind[p], jj = S->ind[p + np], kk = S->ind[p + 2 * np];
                    int ni = i, nj = j, nk = k;
                    if (S->dir == 1) ++ni; else if (S->dir == 2) ++nj; else ++nk;
                    const double jb = (ni <= n && nj <= n && nk <= n) ? S->jmat[(ni + n) + w * ((nj + n) + w * (long)(nk + n))] : 0.0;
                    const double J_av = 0.5 * (S->jmat[p] + jb);
                    if (S->dir == 1) AT(s->Ex, ii, jj, kk) = AT(s->Ex, ii, jj, kk) + c_src * J_av;
                    else if (S->dir == 2) AT(s->Ey, ii, jj, kk) = AT(s->Ey, ii, jj, kk) + c_src * J_av;
                    else AT(s->Ez, ii, jj, kk) = AT(s->Ez, ii, jj, kk) + c_src * J_av;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* detectors: write_3D_field stencils (write_fields_subs_mod.f90:813-1509) */
int orc_add_detector(orc_sim* s, int type, int field, int nd, const int* idx) {
    detector* d = &s->det[s->n_det++];
    d->type = type; d->field = field; d->nd = nd;
    d->idx = (int*)malloc(sizeof(int) * 3 * (nd > 0 ? nd : 1));
    memcpy(d->idx, idx, sizeof(int) * 3 * nd);
    s->n_samples += nd;
    return 0;
}

void orc_sample(const orc_sim* s, double* out) {
    long w = 0;
    for (int q = 0; q < s->n_det; ++q) {
        const detector* d = &s->det[q];
        for (int n = 0; n < d->nd; ++n) {
            const int i = d->idx[n], j = d->idx[n + d->nd], k = d->idx[n + 2 * d->nd];
            double v = 0.0;
            switch (d->field) {
                case 1: v = 0.5 * (AT(s->Ex, i - 1, j, k) + AT(s->Ex, i, j, k)); break;
                case 2: v = 0.5 * (AT(s->Ey, i, j - 1, k) + AT(s->Ey, i, j, k)); break;
                case 3: v = d->type == 1 ? AT(s->Ez, i, j, k) : 0.5 * (AT(s->Ez, i, j, k - 1) + AT(s->Ez, i, j, k)); break;
                case 4: v = 0.25 * (AT(s->Hx, i, j - 1, k) + AT(s->Hx, i, j, k) + AT(s->Hx, i, j, k - 1) + AT(s->Hx, i, j - 1, k - 1)); break;
                case 5: v = 0.25 * (AT(s->Hy, i - 1, j, k) + AT(s->Hy, i, j, k) + AT(s->Hy, i, j, k - 1) + AT(s->Hy, i - 1, j, k - 1)); break;
                case 6: v = 0.25 * (AT(s->Hz, i, j, k) + AT(s->Hz, i - 1, j, k) + AT(s->Hz, i, j - 1, k) + AT(s->Hz, i - 1, j - 1, k)); break;
            }
            out[w++] = v;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* quantum systems in q-groups, non-MPI branches.  group_type: 1 material,
 * 2 single; group_of: each system's group (0-based, systems in group order;
 * NULL for one group); ijk: the system's cell, an (n,3) Fortran array. */
void orc_set_q_groups(orc_sim* s, int n_groups, const int* group_type, int n, const int* group_of, const int* ijk,
                      double density, double dt_q) {
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    s->nq = n;
    s->q_ijk = (int*)malloc(sizeof(int) * 3 * n);
    memcpy(s->q_ijk, ijk, sizeof(int) * 3 * n);
    s->q_ngroups = n_groups;
    s->q_gtype = (int*)malloc(sizeof(int) * n_groups);
    memcpy(s->q_gtype, group_type, sizeof(int) * n_groups);
    s->q_gof = (int*)calloc((size_t)n, sizeof(int));
    if (group_of) memcpy(s->q_gof, group_of, sizeof(int) * n);
    /* "single" kernel, q_group_mod.f90:216-239 (3D): n_ker = 3, r_max =
     * dble(n_ker*2+4), r0 = r_max/2, Blackman-Harris in r = |(i,j,k)|,
     * clipped at 0; the norm it sums is never applied */
    {
        const double aBH[4] = {0.353222222, -0.488, 0.145, -0.010222222};
        const double pi0 = 3.141592653589793, r_max = (double)(3 * 2 + 4), r0 = r_max / 2.0;
        for (int i = -3; i <= 3; ++i)
            for (int j = -3; j <= 3; ++j)
                for (int k = -3; k <= 3; ++k) {
                    const double x = (double)i, y = (double)j, z = (double)k;
                    const double r = sqrt(x * x + y * y + z * z);
                    double v = aBH[0] + aBH[1] * cos(2.0 * pi0 * (r - r0) / r_max) +
                               aBH[2] * cos(2.0 * pi0 * 2.0 * (r - r0) / r_max) +
                               aBH[3] * cos(2.0 * pi0 * 3.0 * (r - r0) / r_max);
                    if (v < 0.0) v = 0.0;
                    s->q_w[(k + 3) * 49 + (j + 3) * 7 + (i + 3)] = v;
                }
    }
    /* mxim_mxll.f90:73-79: n_skip_steps = NINT(dt_q/dt) when dt_q > dt, else 1,
     * and the quantum step becomes n_skip_steps*dt */
    s->n_skip = dt_q > s->dt ? (int)lround(dt_q / s->dt) : 1;
    s->q_density = density; s->q_dt = (double)s->n_skip * s->dt;
    /* J arrays (mxll_3D_mod.f90:132-143) */
    arr_init(&s->Jx, 1, nx + 1, 1, ny, 1, nz); arr_init(&s->Jy, 1, nx, 1, ny + 1, 1, nz); arr_init(&s->Jz, 1, nx, 1, ny, 1, nz + 1);
    arr_init(&s->Jx_old, 1, nx + 1, 1, ny, 1, nz); arr_init(&s->Jy_old, 1, nx, 1, ny + 1, 1, nz); arr_init(&s->Jz_old, 1, nx, 1, ny, 1, nz + 1);
    arr_init(&s->dJx, 1, nx + 1, 1, ny, 1, nz); arr_init(&s->dJy, 1, nx, 1, ny + 1, 1, nz); arr_init(&s->dJz, 1, nx, 1, ny, 1, nz + 1);
    s->have_j = 1;
}

void orc_set_q_sites(orc_sim* s, int n, const int* ijk, double density, double dt_q) {
    const int material = 1;
    orc_set_q_groups(s, 1, &material, n, NULL, ijk, density, dt_q);
}

/* send_E_3D_to_J_ranks, non-MPI branches (interactions_mod.f90:656-691): at
 * the system's cell (a single group's kernel centre) */
void orc_gather_e(const orc_sim* s, double* e) {
    const int n = s->nq;
    for (int q = 0; q < n; ++q) {
        const int i = s->q_ijk[q], j = s->q_ijk[q + n], k = s->q_ijk[q + 2 * n];
        e[q] = 0.5 * (AT(s->Ex, i - 1, j, k) + AT(s->Ex, i, j, k));
        e[q + n] = 0.5 * (AT(s->Ey, i, j - 1, k) + AT(s->Ey, i, j, k));
        e[q + 2 * n] = 0.5 * (AT(s->Ez, i, j, k - 1) + AT(s->Ez, i, j, k));
    }
}

/* send_J_to_E_3D_ranks, non-MPI branches, once per group in group order
 * (interactions_mod.f90:1067-1069,1177-1218, called from mxim_mxll.f90:117-121) */
void orc_scatter_j(orc_sim* s, const double* dpdt) {
    const int n = s->nq;
    for (int g = 0; g < s->q_ngroups; ++g) {
        memcpy(s->Jx_old.p, s->Jx.p, sizeof(double) * s->Jx.n);
        memcpy(s->Jy_old.p, s->Jy.p, sizeof(double) * s->Jy.n);
        memcpy(s->Jz_old.p, s->Jz.p, sizeof(double) * s->Jz.n);
        for (int q = 0; q < n; ++q) {
            if (s->q_gof[q] != g) continue;
            const int i = s->q_ijk[q], j = s->q_ijk[q + n], k = s->q_ijk[q + 2 * n];
            if (s->q_gtype[g] == 1) {
                AT(s->Jx, i, j, k) = s->q_density * dpdt[q];
                AT(s->Jy, i, j, k) = s->q_density * dpdt[q + n];
                AT(s->Jz, i, j, k) = s->q_density * dpdt[q + 2 * n];
                continue;
            }
            for (int kk = -3; kk <= 3; ++kk)
                for (int jj = -3; jj <= 3; ++jj)
                    for (int ii = -3; ii <= 3; ++ii) {
                        const double w = s->q_w[(kk + 3) * 49 + (jj + 3) * 7 + (ii + 3)];
                        AT(s->Jx, i + ii, j + jj, k + kk) = s->q_density * w * dpdt[q];
                        AT(s->Jy, i + ii, j + jj, k + kk) = s->q_density * w * dpdt[q + n];
                        AT(s->Jz, i + ii, j + jj, k + kk) = s->q_density * w * dpdt[q + 2 * n];
                    }
        }
        for (long p = 0; p < s->Jx.n; ++p) s->dJx.p[p] = (s->Jx.p[p] - s->Jx_old.p[p]) / s->q_dt;
        for (long p = 0; p < s->Jy.n; ++p) s->dJy.p[p] = (s->Jy.p[p] - s->Jy_old.p[p]) / s->q_dt;
        for (long p = 0; p < s->Jz.n; ++p) s->dJz.p[p] = (s->Jz.p[p] - s->Jz_old.p[p]) / s->q_dt;
    }
}

/* Two-level emitter (no reference counterpart; TQ_sys_base contract,
 * q_sys_base_mod.f90:7-71): FP64 RK4 over dt_q with the gathered E held
 * fixed; dipole = 2 mu Re(rho_eg) pol; dPt_dt = (dipole - dip_old)/dt_q. */
void orc_set_two_level(orc_sim* s, const double* omega0, const double* mu, const double* pol, const double* gamma2) {
    const int n = s->nq;
    s->tls = (double*)calloc((size_t)16 * n, sizeof(double));
    for (int q = 0; q < n; ++q) {
        double* t = s->tls + 16 * q;
        const double px = pol[q], py = pol[q + n], pz = pol[q + 2 * n];
        const double nr = sqrt(px * px + py * py + pz * pz);
        t[0] = omega0[q]; t[1] = mu[q]; t[2] = px / nr; t[3] = py / nr; t[4] = pz / nr; t[5] = gamma2[q]; t[6] = 1.0;
    }
}

static void tls_rhs(double w0, double V, double g2, const double* y, double* d) {
    d[0] = 2.0 * V * y[3];
    d[1] = -2.0 * V * y[3];
    d[2] = w0 * y[3] - g2 * y[2];
    d[3] = -w0 * y[2] - V * (y[0] - y[1]) - g2 * y[3];
}

void orc_two_level_step(orc_sim* s) {
    const int n = s->nq;
    double* e = (double*)malloc(sizeof(double) * 3 * n);
    double* dpdt = (double*)malloc(sizeof(double) * 3 * n);
    orc_gather_e(s, e);
    for (int q = 0; q < n; ++q) {
        double* t = s->tls + 16 * q;
        const double ep = e[q] * t[2] + e[q + n] * t[3] + e[q + 2 * n] * t[4];
        const double V = -t[1] * ep, w0 = t[0], g2 = t[5], h = s->q_dt;
        double y[4] = {t[6], t[7], t[8], t[9]}, k1[4], k2[4], k3[4], k4[4], yt[4];
        tls_rhs(w0, V, g2, y, k1);
        for (int c = 0; c < 4; ++c) yt[c] = y[c] + 0.5 * h * k1[c];
        tls_rhs(w0, V, g2, yt, k2);
        for (int c = 0; c < 4; ++c) yt[c] = y[c] + 0.5 * h * k2[c];
        tls_rhs(w0, V, g2, yt, k3);
        for (int c = 0; c < 4; ++c) yt[c] = y[c] + h * k3[c];
        tls_rhs(w0, V, g2, yt, k4);
        for (int c = 0; c < 4; ++c) y[c] = y[c] + (h / 6.0) * (k1[c] + 2.0 * k2[c] + 2.0 * k3[c] + k4[c]);
        t[6] = y[0]; t[7] = y[1]; t[8] = y[2]; t[9] = y[3];
        for (int c = 0; c < 3; ++c) {
            const double dn = 2.0 * t[1] * y[2] * t[2 + c];
            t[13 + c] = t[10 + c];
            t[10 + c] = dn;
            dpdt[q + n * c] = (dn - t[13 + c]) / h;
        }
    }
    orc_scatter_j(s, dpdt);
    free(e); free(dpdt);
}

void orc_get_two_level(const orc_sim* s, double* dipole, double* rho) {
    const int n = s->nq;
    for (int q = 0; q < n; ++q) {
        const double* t = s->tls + 16 * q;
        for (int a = 0; a < 3; ++a) dipole[q + n * a] = t[10 + a];
        for (int a = 0; a < 4; ++a) rho[q + n * a] = t[6 + a];
    }
}

/* ------------------------------------------------------------------------ */
/* field access with the reference's array bounds */
void orc_copy_field(orc_sim* s, int field, double* host, int to_sim) {
    arr3* a[6] = {&s->Ex, &s->Ey, &s->Ez, &s->Hx, &s->Hy, &s->Hz};
    if (to_sim) memcpy(a[field - 1]->p, host, sizeof(double) * a[field - 1]->n);
    else memcpy(host, a[field - 1]->p, sizeof(double) * a[field - 1]->n);
}

/* ------------------------------------------------------------------------ */
/* TF/SF plane wave (no reference counterpart), same construction as the engine */
void orc_set_tfsf(orc_sim* s, const int* lo, const int* hi, double e0, double w0, double t0, double tau, int max_steps) {
    for (int a = 0; a < 3; ++a) { s->tf_lo[a] = lo[a]; s->tf_hi[a] = hi[a]; }
    const int pad = max_steps / 2 + 16;
    s->tf_ks = lo[2] - 3;
    s->tf_kb = s->tf_ks - pad;
    s->tf_n = hi[2] + pad - s->tf_kb + 1;
    s->tf_e0 = e0; s->tf_w0 = w0; s->tf_t0 = t0; s->tf_tau = tau;
    s->ex1 = (double*)calloc(s->tf_n, sizeof(double));
    s->hy1 = (double*)calloc(s->tf_n, sizeof(double));
    s->tfsf = 1;
}

/* The incident line: ex1 / hy1 at k = kb .. kb+n-1 (n values each). */
void orc_get_tfsf_line(const orc_sim* s, double* ex1, double* hy1, int* kb, int* n) {
    *kb = s->tf_kb; *n = s->tf_n;
    if (ex1) memcpy(ex1, s->ex1, sizeof(double) * s->tf_n);
    if (hy1) memcpy(hy1, s->hy1, sizeof(double) * s->tf_n);
}

static void tfsf_h(orc_sim* s) {
    const int i0 = s->tf_lo[0], j0 = s->tf_lo[1], k0 = s->tf_lo[2], i1 = s->tf_hi[0], j1 = s->tf_hi[1], k1 = s->tf_hi[2];
    const int kb = s->tf_kb;
    for (int j = j0; j <= j1; ++j) for (int i = i0; i <= i1 - 1; ++i) {
        AT(s->Hy, i, j, k0 - 1) += s->dt_mu0 * s->den_hz[k0 - 1] * s->ex1[k0 - kb];
        AT(s->Hy, i, j, k1) -= s->dt_mu0 * s->den_hz[k1] * s->ex1[k1 - kb];
    }
    for (int k = k0; k <= k1; ++k) for (int i = i0; i <= i1 - 1; ++i) {
        AT(s->Hz, i, j0 - 1, k) -= s->dt_mu0 * s->den_hy[j0 - 1] * s->ex1[k - kb];
        AT(s->Hz, i, j1, k) += s->dt_mu0 * s->den_hy[j1] * s->ex1[k - kb];
    }
    for (int m = 0; m < s->tf_n - 1; ++m) s->hy1[m] = s->hy1[m] + s->dt_mu0 * (1.0 / s->dr) * (s->ex1[m] - s->ex1[m + 1]);
}

static void tfsf_e(orc_sim* s, double t_e) {
    const int i0 = s->tf_lo[0], j0 = s->tf_lo[1], k0 = s->tf_lo[2], i1 = s->tf_hi[0], j1 = s->tf_hi[1], k1 = s->tf_hi[2];
    const int kb = s->tf_kb;
    for (int j = j0; j <= j1; ++j) for (int i = i0; i <= i1 - 1; ++i) {
        AT(s->Ex, i, j, k0) += s->dt_eps0 * s->den_ez[k0] * s->hy1[k0 - 1 - kb];
        AT(s->Ex, i, j, k1) -= s->dt_eps0 * s->den_ez[k1] * s->hy1[k1 - kb];
    }
    for (int k = k0; k <= k1 - 1; ++k) for (int j = j0; j <= j1; ++j) {
        AT(s->Ez, i0, j, k) -= s->dt_eps0 * s->den_ex[i0] * s->hy1[k - kb];
        AT(s->Ez, i1, j, k) += s->dt_eps0 * s->den_ex[i1] * s->hy1[k - kb];
    }
    const double x = (t_e - s->tf_t0) / s->tf_tau;
    const double src = s->tf_e0 * exp(-(x * x)) * cos(s->tf_w0 * (t_e - s->tf_t0));
    for (int m = 1; m < s->tf_n - 1; ++m) {
        double v = s->ex1[m] + s->dt_eps0 * (1.0 / s->dr) * (s->hy1[m - 1] - s->hy1[m]);
        if (m == s->tf_ks - kb) v += src;
        s->ex1[m] = v;
    }
}

/* ------------------------------------------------------------------------ */
/* the driver loop of mxim_mxll.f90:92-132: the quantum systems move, and
 * t_skip is set, on the steps that start a new quantum step,
 * tq_step = 1 + int((tt - tt_0)/n_skip_steps) with tt_0 = 1 when
 * n_skip_steps = 1 and 0 otherwise (mxim_mxll.f90:76-77,97-104) */
long orc_n_samples(const orc_sim* s) { return s->n_samples; }

void orc_run(orc_sim* s, int tt_first, int n_steps, int record_every, double* rec, int* rows) {
    int r = 0;
    const int tt_0 = s->n_skip == 1 ? 1 : 0;
    for (int tt = tt_first; tt < tt_first + n_steps; ++tt) {
        const double time = s->dt * ((double)tt - 0.5);
        const int tq_step = 1 + (tt - tt_0) / s->n_skip;
        int move = 0;
        if (tq_step > s->tq_step_old) {
            s->t_skip = time;
            s->tq_step_old = tq_step;
            move = s->nq > 0;
        }
        orc_step_h(s);
        if (s->tfsf) tfsf_h(s);
        orc_inject_sources(s, time);
        orc_step_e(s, time, s->t_skip);
        if (s->tfsf) tfsf_e(s, s->dt * (double)tt);
        if (s->tls && move) orc_two_level_step(s);
        if (record_every > 0 && tt % record_every == 0 && s->n_samples > 0) {
            if (rec) orc_sample(s, rec + (long)r * s->n_samples);
            ++r;
        }
    }
    if (rows) *rows = r;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================================================================== */
/* 2D and 1D grids: TMxll_2D (mxll_2D_mod.f90) and TMxll_1D (mxll_1D_mod.f90),
 * non-MPI.  Arrays keep the reference's bounds (third index 1..1); the 1D
 * line is stored along the first index (Ex(1:nz), Hy(1:nz-1)), with Hy(0)
 * and Hy(nz) kept at 0 for the Hy detector stencil at the two ends (the
 * reference reads past its array there).  mode: 1 TMz, 2 TEz, 3 both (the
 * reference's FULL mode allocates only the TEz arrays, mxll_2D_mod.f90:110-163,
 * and cannot run; here both families run, each as in its own mode). */
typedef struct {
    int dir, n_ker, npts;
    double j_amp, w0, t0, tau, t_init, t_final, phase;
    int *ind;        /* npts*2 (1D: npts) */
    double *ker, *jmat;
} lsource;

typedef struct orl_sim {
    int dims, mode, nx, ny, npml, bc[2], cpml_pos[4];
    double dt, dr, dt_eps0, dt_mu0;
    arr3 Ex, Ey, Ez, Hx, Hy, Hz, Ex_old, Ey_old, Ez_old, eps_x, eps_y, eps_z, PDx, PDy, PDz;
    double *PLx, *PLy, *PLz, *PLx_old, *PLy_old, *PLz_old;
    arr3 psix_Ezx, psix_Eyx, psiy_Exy, psiy_Ezy, psix_Hyx, psix_Hzx, psiy_Hxy, psiy_Hzy;
    double *den_ex, *den_ey, *den_hx, *den_hy, *be, *ce, *bh, *ch;
    int *media_map, n_media, max_poles;
    medium_co media[16];
    int n_src; lsource src[16];
    int n_det; int det_field[64], det_nd[64]; int* det_idx[64]; long n_samples;
} orl_sim;

/* CPML profiles of init_2Dgrid / init_1Dgrid (mxll_2D_mod.f90:270-300,
 * mxll_1D_mod.f90:150-176), the same formulas as the 3D grid. */
static void orl_profiles(orl_sim* s, double* kappae, double* kappah) {
    const int npml = s->npml;
    const double sec_to_au = 1.0 / 2.4188843265864e-17, C_to_au = 1.0 / 1.602176634e-19;
    const double kg_to_au = 1.0 / 9.1093837139e-31, m_to_au = 1.0 / 5.29177210544e-11;
    const double sig_to_au = C_to_au * C_to_au * sec_to_au / (kg_to_au * (m_to_au * m_to_au * m_to_au));
    const double sigmaCPML = (double)(0.8f * (float)(3 + 1)) / (s->dr * sqrt(MU0 / EPS0));
    double sig, al, kap;
    for (int kk = 1; kk <= npml; ++kk) {
        grade_e(npml, sigmaCPML, sig_to_au, kk, &sig, &al, &kap);
        kappae[kk] = kap;
        s->be[kk] = exp(-(sig / kap + al) * s->dt / EPS0);
        if (sig == 0.0 && al == 0.0 && kk == npml) s->ce[kk] = 0.0;
        else s->ce[kk] = sig * (s->be[kk] - 1.0) / (sig + kap * al) / kap;
    }
    for (int kk = 1; kk <= npml - 1; ++kk) {
        grade_h(npml, sigmaCPML, sig_to_au, kk, &sig, &al, &kap);
        kappah[kk] = kap;
        s->bh[kk] = exp(-(sig / kap + al) * s->dt / EPS0);
        s->ch[kk] = sig * (s->bh[kk] - 1.0) / (sig + kap * al) / kap;
    }
}

/* init_2Dgrid (mxll_2D_mod.f90:177-420) / init_1Dgrid (mxll_1D_mod.f90:52-210). */
orl_sim* orl_create(int dims, int mode, int n1, int n2, int npml, const int* bc, double dt, double dr) {
    orl_sim* s = (orl_sim*)calloc(1, sizeof(orl_sim));
    s->dims = dims; s->mode = dims == 1 ? 0 : mode; s->npml = npml;
    s->nx = n1; s->ny = dims == 1 ? 1 : n2;
    s->bc[0] = bc[0]; s->bc[1] = dims == 1 ? 1 : bc[1];
    s->dt = dt; s->dr = dr; s->dt_eps0 = dt / EPS0; s->dt_mu0 = dt / MU0;
    const int nx = s->nx, ny = s->ny;
    s->den_ex = (double*)calloc(nx + 2, sizeof(double)); s->den_hx = (double*)calloc(nx + 2, sizeof(double));
    s->den_ey = (double*)calloc(ny + 2, sizeof(double)); s->den_hy = (double*)calloc(ny + 2, sizeof(double));
    for (int k = 1; k <= nx; ++k) s->den_ex[k] = s->den_hx[k] = 1.0 / dr;
    for (int k = 1; k <= ny; ++k) s->den_ey[k] = s->den_hy[k] = 1.0 / dr;
    const int np = npml > 2 ? npml : 2;
    s->be = (double*)calloc(np + 1, sizeof(double)); s->ce = (double*)calloc(np + 1, sizeof(double));
    s->bh = (double*)calloc(np + 1, sizeof(double)); s->ch = (double*)calloc(np + 1, sizeof(double));
    double* kappae = (double*)calloc(np + 1, sizeof(double));
    double* kappah = (double*)calloc(np + 1, sizeof(double));
    if (dims == 1) {
        const int nz = nx;
        arr_init(&s->Ex, 1, nz, 1, 1, 1, 1); arr_init(&s->Ex_old, 1, nz, 1, 1, 1, 1);
        arr_init(&s->Hy, 0, nz, 1, 1, 1, 1);
        arr_init(&s->eps_x, 1, nz, 1, 1, 1, 1);
        for (long q = 0; q < s->eps_x.n; ++q) s->eps_x.p[q] = 1.0;
        arr_init(&s->PDx, 1, nz, 1, 1, 1, 1);
        s->media_map = (int*)calloc((size_t)nz + 1, sizeof(int));
        if (bc[0] == 3) {
            s->cpml_pos[0] = s->cpml_pos[1] = 1;
            arr_init(&s->psix_Ezx, 1, 2 * npml, 1, 1, 1, 1);        /* psiz_Exz */
            arr_init(&s->psix_Hyx, 1, 2 * (npml - 1), 1, 1, 1, 1);  /* psiz_Hyz */
            orl_profiles(s, kappae, kappah);
            int kk = npml;
            for (int k = 1; k <= nz - 1; ++k) {
                if (k <= npml) s->den_ex[k] = 1.0 / (kappae[k] * dr);
                else if (k >= nz + 1 - npml) { s->den_ex[k] = 1.0 / (kappae[kk] * dr); kk = kk - 1; }
                else s->den_ex[k] = 1.0 / dr;
            }
            kk = npml - 1;
            for (int k = 1; k <= nz - 1; ++k) {
                if (k <= npml - 1) s->den_hx[k] = 1.0 / (kappah[k] * dr);
                else if (k >= nz + 1 - npml) { s->den_hx[k] = 1.0 / (kappah[kk] * dr); kk = kk - 1; }
                else s->den_hx[k] = 1.0 / dr;
            }
        }
    } else {
        const int te = mode & 2, tm = mode & 1;
        if (te) {
            arr_init(&s->Ex, 0, nx, 1, ny + 1, 1, 1); arr_init(&s->Ex_old, 0, nx, 1, ny + 1, 1, 1);
            arr_init(&s->Ey, 1, nx + 1, 0, ny, 1, 1); arr_init(&s->Ey_old, 1, nx + 1, 0, ny, 1, 1);
            arr_init(&s->Hz, 0, nx, 0, ny, 1, 1);
        }
        if (tm) {
            arr_init(&s->Ez, 1, nx + 1, 1, ny + 1, 1, 1); arr_init(&s->Ez_old, 1, nx + 1, 1, ny + 1, 1, 1);
            arr_init(&s->Hx, 1, nx, 0, ny, 1, 1);
            arr_init(&s->Hy, 0, nx, 1, ny, 1, 1);
        }
        arr_init(&s->eps_x, 1, nx, 1, ny, 1, 1); arr_init(&s->eps_y, 1, nx, 1, ny, 1, 1); arr_init(&s->eps_z, 1, nx, 1, ny, 1, 1);
        for (long q = 0; q < s->eps_x.n; ++q) s->eps_x.p[q] = s->eps_y.p[q] = s->eps_z.p[q] = 1.0;
        arr_init(&s->PDx, 1, nx, 1, ny, 1, 1); arr_init(&s->PDy, 1, nx, 1, ny, 1, 1); arr_init(&s->PDz, 1, nx, 1, ny, 1, 1);
        s->media_map = (int*)calloc((size_t)nx * ny * 3, sizeof(int));
        s->cpml_pos[0] = s->cpml_pos[1] = bc[0] == 3;
        s->cpml_pos[2] = s->cpml_pos[3] = bc[1] == 3;
        const int n_sec = 2;
        if (s->cpml_pos[0]) {
            if (te) { arr_init(&s->psix_Eyx, 1, n_sec * npml, 1, ny, 1, 1); arr_init(&s->psix_Hzx, 1, n_sec * (npml - 1), 1, ny, 1, 1); }
            if (tm) { arr_init(&s->psix_Hyx, 1, n_sec * (npml - 1), 1, ny, 1, 1); arr_init(&s->psix_Ezx, 1, n_sec * npml, 1, ny, 1, 1); }
        }
        if (s->cpml_pos[2]) {
            if (te) { arr_init(&s->psiy_Exy, 1, nx, 1, n_sec * npml, 1, 1); arr_init(&s->psiy_Hzy, 1, nx, 1, n_sec * (npml - 1), 1, 1); }
            if (tm) { arr_init(&s->psiy_Ezy, 1, nx, 1, n_sec * npml, 1, 1); arr_init(&s->psiy_Hxy, 1, nx, 1, n_sec * (npml - 1), 1, 1); }
        }
        if (s->cpml_pos[0] || s->cpml_pos[2]) {
            orl_profiles(s, kappae, kappah);
            const int n[2] = {nx, ny};
            double* de[2] = {s->den_ex, s->den_ey};
            double* dh[2] = {s->den_hx, s->den_hy};
            for (int a = 0; a < 2; ++a) {
                if (s->cpml_pos[2 * a]) {
                    for (int k = 1; k <= n[a]; ++k) if (k <= npml) de[a][k] = 1.0 / (kappae[k] * dr);
                    for (int k = 1; k <= n[a]; ++k) if (k <= npml - 1) dh[a][k] = 1.0 / (kappah[k] * dr);
                }
                if (s->cpml_pos[2 * a + 1]) {
                    int kk = npml;
                    for (int k = 1; k <= n[a] - 1; ++k) if (k >= n[a] + 1 - npml) { de[a][k] = 1.0 / (kappae[kk] * dr); kk = kk - 1; }
                    kk = npml - 1;
                    for (int k = 1; k <= n[a] - 1; ++k) if (k >= n[a] + 1 - npml) { dh[a][k] = 1.0 / (kappah[kk] * dr); kk = kk - 1; }
                }
            }
        }
    }
    free(kappae); free(kappah);
    s->max_poles = 1;
    return s;
}

void orl_destroy(orl_sim* s) {
    if (!s) return;
    arr3* arrs[] = {&s->Ex, &s->Ey, &s->Ez, &s->Hx, &s->Hy, &s->Hz, &s->Ex_old, &s->Ey_old, &s->Ez_old, &s->eps_x,
                    &s->eps_y, &s->eps_z, &s->PDx, &s->PDy, &s->PDz, &s->psix_Ezx, &s->psix_Eyx, &s->psiy_Exy,
                    &s->psiy_Ezy, &s->psix_Hyx, &s->psix_Hzx, &s->psiy_Hxy, &s->psiy_Hzy};
    for (size_t q = 0; q < sizeof(arrs) / sizeof(arrs[0]); ++q) free(arrs[q]->p);
    free(s->PLx); free(s->PLy); free(s->PLz); free(s->PLx_old); free(s->PLy_old); free(s->PLz_old);
    free(s->den_ex); free(s->den_ey); free(s->den_hx); free(s->den_hy);
    free(s->be); free(s->ce); free(s->bh); free(s->ch); free(s->media_map);
    for (int q = 0; q < s->n_src; ++q) { free(s->src[q].ind); free(s->src[q].ker); free(s->src[q].jmat); }
    for (int q = 0; q < s->n_det; ++q) free(s->det_idx[q]);
    free(s);
}

void orl_get_den(const orl_sim* s, double* out) {   /* den_ex(1:n1) den_hx den_ey(1:n2) den_hy */
    long w = 0;
    for (int k = 1; k <= s->nx; ++k) out[w++] = s->den_ex[k];
    for (int k = 1; k <= s->nx; ++k) out[w++] = s->den_hx[k];
    for (int k = 1; k <= s->ny; ++k) out[w++] = s->den_ey[k];
    for (int k = 1; k <= s->ny; ++k) out[w++] = s->den_hy[k];
}

/* read_init_media_2D after the point lists are read (classical_medium_mod.f90:484-538)
 * and read_init_media_1D (:203-337); aux is the cell medium map. */
int orl_set_media(orl_sim* s, int n_media, const orc_medium* table, const int8_t* aux) {
    const int nx = s->nx, ny = s->ny;
    s->n_media = n_media;
    for (int n = 1; n <= n_media; ++n) {
        medium_coeffs(&table[n - 1], s->dt, &s->media[n]);
        if (table[n - 1].kind == 3 && table[n - 1].n_poles > s->max_poles) s->max_poles = table[n - 1].n_poles;
    }
    if (s->dims == 1) {
        /* dielectric cells: eps_x = 1/eps_r; the map takes each cell's medium */
        for (int i = 1; i <= nx; ++i) {
            const int n = aux[i - 1];
            s->media_map[i] = n;
            if (n && s->media[n].type == 1) AT(s->eps_x, i, 1, 1) = 1.0 / s->media[n].eps_r;
        }
    } else {
#define AUX2(i, j) (((i) > nx || (j) > ny) ? 0 : (int)aux[((i) - 1) + (long)nx * ((j) - 1)])
#define MAP2(i, j, c) s->media_map[((i) - 1) + (long)nx * (((j) - 1) + (long)ny * (c))]
        for (int n = 1; n <= n_media; ++n)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    if (i == nx && AUX2(i, j) == n) MAP2(i, j, 0) = n;
                    else if (j == ny && AUX2(i, j) == n) MAP2(i, j, 0) = n;
                    else {
                        if (AUX2(i, j) == n && AUX2(i + 1, j) == n) MAP2(i, j, 0) = n;
                        else if (AUX2(i, j) != 0 && AUX2(i + 1, j) == n) MAP2(i, j, 0) = n;
                        if (AUX2(i, j) == n && AUX2(i, j + 1) == n) MAP2(i, j, 1) = n;
                        else if (AUX2(i, j) != 0 && AUX2(i, j + 1) == n) MAP2(i, j, 1) = n;
                        if (AUX2(i, j) == n) MAP2(i, j, 2) = n;
                    }
                }
        for (int n = 1; n <= n_media; ++n) {
            if (s->media[n].type != 1) continue;
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i)
                    if (MAP2(i, j, 0) == n) {
                        AT(s->eps_x, i, j, 1) = 1.0 / s->media[n].eps_r;
                        AT(s->eps_y, i, j, 1) = 1.0 / s->media[n].eps_r;
                        AT(s->eps_z, i, j, 1) = 1.0 / s->media[n].eps_r;
                    }
        }
#undef AUX2
    }
    const size_t np = (size_t)nx * ny * s->max_poles;
    s->PLx = (double*)calloc(np, sizeof(double)); s->PLy = (double*)calloc(np, sizeof(double));
    s->PLz = (double*)calloc(np, sizeof(double)); s->PLx_old = (double*)calloc(np, sizeof(double));
    s->PLy_old = (double*)calloc(np, sizeof(double)); s->PLz_old = (double*)calloc(np, sizeof(double));
    return 0;
}

void orl_get_media_map(const orl_sim* s, int32_t* out) {
    if (s->dims == 1) memcpy(out, s->media_map + 1, sizeof(int) * (size_t)s->nx);
    else memcpy(out, s->media_map, sizeof(int) * (size_t)s->nx * s->ny * 3);
}

/* td_propagate_H_2Dfield (mxll_2D_mod.f90:450-636) / td_propagate_H_1Dfield
 * (mxll_1D_mod.f90:252-293) */
void orl_step_h(orl_sim* s) {
    const int nx = s->nx, ny = s->ny, npml = s->npml;
    if (s->dims == 1) {
        const int nz = nx;
        /* periodic: Ex(nz+1) = Ex(1) lands past Ex(1:nz) and is never read */
        for (int i = 1; i <= nz - 1; ++i)
            AT(s->Hy, i, 1, 1) = AT(s->Hy, i, 1, 1) + s->dt_mu0 * (AT(s->Ex, i, 1, 1) - AT(s->Ex, i + 1, 1, 1)) * s->den_hx[i];
        if (s->bc[0] == 3) {
            for (int i = 1; i <= npml - 1; ++i) {
                AT(s->psix_Hyx, i, 1, 1) = s->bh[i] * AT(s->psix_Hyx, i, 1, 1) + s->ch[i] * (AT(s->Ex, i, 1, 1) - AT(s->Ex, i + 1, 1, 1)) / s->dr;
                AT(s->Hy, i, 1, 1) = AT(s->Hy, i, 1, 1) + s->dt_mu0 * AT(s->psix_Hyx, i, 1, 1);
            }
            int kk = 2 * (npml - 1);
            for (int i = nz + 1 - npml; i <= nz - 1; ++i) {
                AT(s->psix_Hyx, kk, 1, 1) = s->bh[kk - (npml - 1)] * AT(s->psix_Hyx, kk, 1, 1) +
                                            s->ch[kk - (npml - 1)] * (AT(s->Ex, i, 1, 1) - AT(s->Ex, i + 1, 1, 1)) / s->dr;
                AT(s->Hy, i, 1, 1) = AT(s->Hy, i, 1, 1) + s->dt_mu0 * AT(s->psix_Hyx, kk, 1, 1);
                kk = kk - 1;
            }
        }
        return;
    }
    const int n_sec = 2, n_base = npml - 1;
    if (s->mode & 1) {
        if (s->bc[0] == 2) for (int j = 1; j <= ny + 1; ++j) AT(s->Ez, nx + 1, j, 1) = AT(s->Ez, 1, j, 1);
        if (s->bc[1] == 2) for (int i = 1; i <= nx + 1; ++i) AT(s->Ez, i, ny + 1, 1) = AT(s->Ez, i, 1, 1);
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny; ++j) {
                const double rotE = (AT(s->Ez, i, j, 1) - AT(s->Ez, i, j + 1, 1)) * s->den_hy[j];
                AT(s->Hx, i, j, 1) = AT(s->Hx, i, j, 1) + s->dt_mu0 * rotE;
            }
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny; ++j) {
                const double rotE = (AT(s->Ez, i + 1, j, 1) - AT(s->Ez, i, j, 1)) * s->den_hx[i];
                AT(s->Hy, i, j, 1) = AT(s->Hy, i, j, 1) + s->dt_mu0 * rotE;
            }
        if (s->cpml_pos[0])
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= npml - 1; ++i) {
                    const double rotE = (AT(s->Ez, i + 1, j, 1) - AT(s->Ez, i, j, 1)) / s->dr;
                    AT(s->psix_Hyx, i, j, 1) = s->bh[i] * AT(s->psix_Hyx, i, j, 1) + s->ch[i] * rotE;
                    AT(s->Hy, i, j, 1) = AT(s->Hy, i, j, 1) + s->dt_mu0 * AT(s->psix_Hyx, i, j, 1);
                }
        if (s->cpml_pos[1])
            for (int j = 1; j <= ny; ++j) {
                int ii = n_sec * (npml - 1);
                for (int i = nx + 1 - npml; i <= nx - 1; ++i) {
                    const double rotE = (AT(s->Ez, i + 1, j, 1) - AT(s->Ez, i, j, 1)) / s->dr;
                    AT(s->psix_Hyx, ii, j, 1) = s->bh[ii - n_base] * AT(s->psix_Hyx, ii, j, 1) + s->ch[ii - n_base] * rotE;
                    AT(s->Hy, i, j, 1) = AT(s->Hy, i, j, 1) + s->dt_mu0 * AT(s->psix_Hyx, ii, j, 1);
                    ii = ii - 1;
                }
            }
        if (s->cpml_pos[2])
            for (int i = 1; i <= nx; ++i)
                for (int j = 1; j <= npml - 1; ++j) {
                    const double rotE = (AT(s->Ez, i, j, 1) - AT(s->Ez, i, j + 1, 1)) / s->dr;
                    AT(s->psiy_Hxy, i, j, 1) = s->bh[j] * AT(s->psiy_Hxy, i, j, 1) + s->ch[j] * rotE;
                    AT(s->Hx, i, j, 1) = AT(s->Hx, i, j, 1) + s->dt_mu0 * AT(s->psiy_Hxy, i, j, 1);
                }
        if (s->cpml_pos[3])
            for (int i = 1; i <= nx; ++i) {
                int jj = n_sec * (npml - 1);
                for (int j = ny + 1 - npml; j <= ny - 1; ++j) {
                    const double rotE = (AT(s->Ez, i, j, 1) - AT(s->Ez, i, j + 1, 1)) / s->dr;
                    AT(s->psiy_Hxy, i, jj, 1) = s->bh[jj - n_base] * AT(s->psiy_Hxy, i, jj, 1) + s->ch[jj - n_base] * rotE;
                    AT(s->Hx, i, j, 1) = AT(s->Hx, i, j, 1) + s->dt_mu0 * AT(s->psiy_Hxy, i, jj, 1);
                    jj = jj - 1;
                }
            }
    }
    if (s->mode & 2) {
        if (s->bc[0] == 2) for (int j = 0; j <= ny; ++j) AT(s->Ey, nx + 1, j, 1) = AT(s->Ey, 1, j, 1);
        if (s->bc[1] == 2) for (int i = 0; i <= nx; ++i) AT(s->Ex, i, ny + 1, 1) = AT(s->Ex, i, 1, 1);
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny; ++j) {
                const double rotE = (AT(s->Ey, i, j, 1) - AT(s->Ey, i + 1, j, 1)) * s->den_hx[i] +
                                    (AT(s->Ex, i, j + 1, 1) - AT(s->Ex, i, j, 1)) * s->den_hy[j];
                AT(s->Hz, i, j, 1) = AT(s->Hz, i, j, 1) + s->dt_mu0 * rotE;
            }
        if (s->cpml_pos[0])
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= npml - 1; ++i) {
                    const double rotE = (AT(s->Ey, i, j, 1) - AT(s->Ey, i + 1, j, 1)) / s->dr;
                    AT(s->psix_Hzx, i, j, 1) = s->bh[i] * AT(s->psix_Hzx, i, j, 1) + s->ch[i] * rotE;
                    AT(s->Hz, i, j, 1) = AT(s->Hz, i, j, 1) + s->dt_mu0 * AT(s->psix_Hzx, i, j, 1);
                }
        if (s->cpml_pos[1])
            for (int j = 1; j <= ny; ++j) {
                int ii = n_sec * (npml - 1);
                for (int i = nx + 1 - npml; i <= nx - 1; ++i) {
                    const double rotE = (AT(s->Ey, i, j, 1) - AT(s->Ey, i + 1, j, 1)) / s->dr;
                    AT(s->psix_Hzx, ii, j, 1) = s->bh[ii - n_base] * AT(s->psix_Hzx, ii, j, 1) + s->ch[ii - n_base] * rotE;
                    AT(s->Hz, i, j, 1) = AT(s->Hz, i, j, 1) + s->dt_mu0 * AT(s->psix_Hzx, ii, j, 1);
                    ii = ii - 1;
                }
            }
        if (s->cpml_pos[2])
            for (int i = 1; i <= nx; ++i)
                for (int j = 1; j <= npml - 1; ++j) {
                    const double rotE = (AT(s->Ex, i, j + 1, 1) - AT(s->Ex, i, j, 1)) / s->dr;
                    AT(s->psiy_Hzy, i, j, 1) = s->bh[j] * AT(s->psiy_Hzy, i, j, 1) + s->ch[j] * rotE;
                    AT(s->Hz, i, j, 1) = AT(s->Hz, i, j, 1) + s->dt_mu0 * AT(s->psiy_Hzy, i, j, 1);
                }
        if (s->cpml_pos[3])
            for (int i = 1; i <= nx; ++i) {
                int jj = n_sec * (npml - 1);
                for (int j = ny + 1 - npml; j <= ny - 1; ++j) {
                    const double rotE = (AT(s->Ex, i, j + 1, 1) - AT(s->Ex, i, j, 1)) / s->dr;
                    AT(s->psiy_Hzy, i, jj, 1) = s->bh[jj - n_base] * AT(s->psiy_Hzy, i, jj, 1) + s->ch[jj - n_base] * rotE;
                    AT(s->Hz, i, j, 1) = AT(s->Hz, i, j, 1) + s->dt_mu0 * AT(s->psiy_Hzy, i, jj, 1);
                    jj = jj - 1;
                }
            }
    }
}

/* one E component of a cell: get_medium_polarization, else the dielectric /
 * vacuum update with J = 0 (no quantum groups on these grids) */
static double orl_e_cell(const orl_sim* s, int m, arr3* PD, double* PL, double* PLo, arr3* Eo, double e, double eps,
                         double rotH, int i, int j) {
    const long c = (i - 1) + (long)s->nx * (j - 1);
    const long ps = (long)s->nx * s->ny;
    double E = e;
    if (medium_polarization(s->media, m, &AT(*PD, i, j, 1), PL + c, PLo + c, ps, &E, &AT(*Eo, i, j, 1), rotH))
        E = E + s->dt_eps0 * eps * rotH - s->dt_eps0 * 0.0;
    return E;
}

/* td_propagate_E_2Dfield (mxll_2D_mod.f90:638-864) / td_propagate_E_1Dfield
 * (mxll_1D_mod.f90:296-350) */
void orl_step_e(orl_sim* s) {
    const int nx = s->nx, ny = s->ny, npml = s->npml;
    if (s->dims == 1) {
        const int nz = nx;
        for (int i = 2; i <= nz - 1; ++i) {
            const double rotH = (AT(s->Hy, i - 1, 1, 1) - AT(s->Hy, i, 1, 1)) * s->den_ex[i];
            AT(s->Ex, i, 1, 1) = orl_e_cell(s, s->media_map[i], &s->PDx, s->PLx, s->PLx_old, &s->Ex_old, AT(s->Ex, i, 1, 1),
                                            AT(s->eps_x, i, 1, 1), rotH, i, 1);
        }
        if (s->bc[0] == 3) {
            for (int i = 2; i <= npml; ++i) {
                const double rotH = (AT(s->Hy, i - 1, 1, 1) - AT(s->Hy, i, 1, 1)) / s->dr;
                AT(s->psix_Ezx, i, 1, 1) = s->be[i] * AT(s->psix_Ezx, i, 1, 1) + s->ce[i] * rotH;
                AT(s->Ex, i, 1, 1) = AT(s->Ex, i, 1, 1) + s->dt_eps0 * AT(s->psix_Ezx, i, 1, 1);
            }
            int kk = 2 * npml;
            for (int i = nz + 1 - npml; i <= nz - 1; ++i) {
                const double rotH = (AT(s->Hy, i - 1, 1, 1) - AT(s->Hy, i, 1, 1)) / s->dr;
                AT(s->psix_Ezx, kk, 1, 1) = s->be[kk - npml] * AT(s->psix_Ezx, kk, 1, 1) + s->ce[kk - npml] * rotH;
                AT(s->Ex, i, 1, 1) = AT(s->Ex, i, 1, 1) + s->dt_eps0 * AT(s->psix_Ezx, kk, 1, 1);
                kk = kk - 1;
            }
        }
        return;
    }
    const int n_sec = 2, n_base = npml;
    const long mstride = (long)nx * ny;
#define MM(i, j, c) s->media_map[((i) - 1) + (long)nx * ((j) - 1) + mstride * (c)]
    if (s->mode & 1) {
        if (s->bc[0] == 2) for (int j = 1; j <= ny; ++j) AT(s->Hy, 0, j, 1) = AT(s->Hy, nx, j, 1);
        if (s->bc[1] == 2) for (int i = 1; i <= nx; ++i) AT(s->Hx, i, 0, 1) = AT(s->Hx, i, ny, 1);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double rotH = (AT(s->Hy, i, j, 1) - AT(s->Hy, i - 1, j, 1)) * s->den_ex[i] +
                                    (AT(s->Hx, i, j - 1, 1) - AT(s->Hx, i, j, 1)) * s->den_ey[j];
                AT(s->Ez, i, j, 1) = orl_e_cell(s, MM(i, j, 2), &s->PDz, s->PLz, s->PLz_old, &s->Ez_old, AT(s->Ez, i, j, 1),
                                                AT(s->eps_z, i, j, 1), rotH, i, j);
            }
        if (s->cpml_pos[0])
            for (int j = 1; j <= ny; ++j)
                for (int i = 2; i <= npml; ++i) {
                    const double rotH = (AT(s->Hy, i, j, 1) - AT(s->Hy, i - 1, j, 1)) / s->dr;
                    AT(s->psix_Ezx, i, j, 1) = s->be[i] * AT(s->psix_Ezx, i, j, 1) + s->ce[i] * rotH;
                    AT(s->Ez, i, j, 1) = AT(s->Ez, i, j, 1) + s->dt_eps0 * AT(s->psix_Ezx, i, j, 1);
                }
        if (s->cpml_pos[1])
            for (int j = 1; j <= ny; ++j) {
                int ii = n_sec * npml;
                for (int i = nx + 1 - npml; i <= nx - 1; ++i) {
                    const double rotH = (AT(s->Hy, i, j, 1) - AT(s->Hy, i - 1, j, 1)) / s->dr;
                    AT(s->psix_Ezx, ii, j, 1) = s->be[ii - n_base] * AT(s->psix_Ezx, ii, j, 1) + s->ce[ii - n_base] * rotH;
                    AT(s->Ez, i, j, 1) = AT(s->Ez, i, j, 1) + s->dt_eps0 * AT(s->psix_Ezx, ii, j, 1);
                    ii = ii - 1;
                }
            }
        if (s->cpml_pos[2])
            for (int i = 1; i <= nx; ++i)
                for (int j = 2; j <= npml; ++j) {
                    const double rotH = (AT(s->Hx, i, j - 1, 1) - AT(s->Hx, i, j, 1)) / s->dr;
                    AT(s->psiy_Ezy, i, j, 1) = s->be[j] * AT(s->psiy_Ezy, i, j, 1) + s->ce[j] * rotH;
                    AT(s->Ez, i, j, 1) = AT(s->Ez, i, j, 1) + s->dt_eps0 * AT(s->psiy_Ezy, i, j, 1);
                }
        if (s->cpml_pos[3])
            for (int i = 1; i <= nx; ++i) {
                int jj = n_sec * npml;
                for (int j = ny + 1 - npml; j <= ny - 1; ++j) {
                    const double rotH = (AT(s->Hx, i, j - 1, 1) - AT(s->Hx, i, j, 1)) / s->dr;
                    AT(s->psiy_Ezy, i, jj, 1) = s->be[jj - n_base] * AT(s->psiy_Ezy, i, jj, 1) + s->ce[jj - n_base] * rotH;
                    AT(s->Ez, i, j, 1) = AT(s->Ez, i, j, 1) + s->dt_eps0 * AT(s->psiy_Ezy, i, jj, 1);
                    jj = jj - 1;
                }
            }
    }
    if (s->mode & 2) {
        if (s->bc[0] == 2) for (int j = 0; j <= ny; ++j) AT(s->Hz, 0, j, 1) = AT(s->Hz, nx, j, 1);
        if (s->bc[1] == 2) for (int i = 0; i <= nx; ++i) AT(s->Hz, i, 0, 1) = AT(s->Hz, i, ny, 1);
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double rotH = (AT(s->Hz, i, j, 1) - AT(s->Hz, i, j - 1, 1)) * s->den_ey[j];
                AT(s->Ex, i, j, 1) = orl_e_cell(s, MM(i, j, 0), &s->PDx, s->PLx, s->PLx_old, &s->Ex_old, AT(s->Ex, i, j, 1),
                                                AT(s->eps_x, i, j, 1), rotH, i, j);
            }
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const double rotH = (AT(s->Hz, i - 1, j, 1) - AT(s->Hz, i, j, 1)) * s->den_ex[i];
                AT(s->Ey, i, j, 1) = orl_e_cell(s, MM(i, j, 1), &s->PDy, s->PLy, s->PLy_old, &s->Ey_old, AT(s->Ey, i, j, 1),
                                                AT(s->eps_y, i, j, 1), rotH, i, j);
            }
        if (s->cpml_pos[0])
            for (int j = 1; j <= ny; ++j)
                for (int i = 2; i <= npml; ++i) {
                    const double rotH = (AT(s->Hz, i - 1, j, 1) - AT(s->Hz, i, j, 1)) / s->dr;
                    AT(s->psix_Eyx, i, j, 1) = s->be[i] * AT(s->psix_Eyx, i, j, 1) + s->ce[i] * rotH;
                    AT(s->Ey, i, j, 1) = AT(s->Ey, i, j, 1) + s->dt_eps0 * AT(s->psix_Eyx, i, j, 1);
                }
        if (s->cpml_pos[1])
            for (int j = 1; j <= ny; ++j) {
                int ii = n_sec * npml;
                for (int i = nx + 1 - npml; i <= nx - 1; ++i) {
                    const double rotH = (AT(s->Hz, i - 1, j, 1) - AT(s->Hz, i, j, 1)) / s->dr;
                    AT(s->psix_Eyx, ii, j, 1) = s->be[ii - n_base] * AT(s->psix_Eyx, ii, j, 1) + s->ce[ii - n_base] * rotH;
                    AT(s->Ey, i, j, 1) = AT(s->Ey, i, j, 1) + s->dt_eps0 * AT(s->psix_Eyx, ii, j, 1);
                    ii = ii - 1;
                }
            }
        if (s->cpml_pos[2])
            for (int i = 1; i <= nx; ++i)
                for (int j = 2; j <= npml; ++j) {
                    const double rotH = (AT(s->Hz, i, j, 1) - AT(s->Hz, i, j - 1, 1)) / s->dr;
                    AT(s->psiy_Exy, i, j, 1) = s->be[j] * AT(s->psiy_Exy, i, j, 1) + s->ce[j] * rotH;
                    AT(s->Ex, i, j, 1) = AT(s->Ex, i, j, 1) + s->dt_eps0 * AT(s->psiy_Exy, i, j, 1);
                }
        if (s->cpml_pos[3])
            for (int i = 1; i <= nx; ++i) {
                int jj = 2 * npml;
                for (int j = ny + 1 - npml; j <= ny - 1; ++j) {
                    const double rotH = (AT(s->Hz, i, j, 1) - AT(s->Hz, i, j - 1, 1)) / s->dr;
                    AT(s->psiy_Exy, i, jj, 1) = s->be[jj - n_base] * AT(s->psiy_Exy, i, jj, 1) + s->ce[jj - n_base] * rotH;
                    AT(s->Ex, i, j, 1) = AT(s->Ex, i, j, 1) + s->dt_eps0 * AT(s->psiy_Exy, i, jj, 1);
                    jj = jj - 1;
                }
            }
    }
#undef MM
}

/* sources: compute_kernel 1D / 2D (source_mod.f90:251-300) and
 * determine_indx_and_ranks (:340-346) */
void orl_source_kernel(int dims, int n_ker, double dr, double* ker) {
    const double aBH[4] = {0.353222222, -0.488, 0.145, -0.010222222};
    const double r_max = (n_ker * 2 + 2) * dr, r0 = r_max / 2.0;
    const int w = 2 * n_ker + 1;
    double norm = 0.0;
    if (dims == 1) {
        for (int i = -n_ker; i <= n_ker; ++i) {
            const double x = i * dr;
            const double v = aBH[0] + aBH[1] * cos(2.0 * pi0 * (x - r0) / r_max) + aBH[2] * cos(2.0 * pi0 * 2.0 * (x - r0) / r_max) +
                             aBH[3] * cos(2.0 * pi0 * 3.0 * (x - r0) / r_max);
            ker[i + n_ker] = v;
            norm = norm + v;
        }
        for (int q = 0; q < w; ++q) ker[q] = ker[q] / norm;
        return;
    }
    for (int j = -n_ker; j <= n_ker; ++j)
        for (int i = -n_ker; i <= n_ker; ++i) {
            const double x = i * dr, y = j * dr;
            const double r = sqrt(x * x + y * y);
            double v = aBH[0] + aBH[1] * cos(2.0 * pi0 * (r - r0) / r_max) + aBH[2] * cos(2.0 * pi0 * 2.0 * (r - r0) / r_max) +
                       aBH[3] * cos(2.0 * pi0 * 3.0 * (r - r0) / r_max);
            if (v < 0.0) v = 0.0;
            ker[(i + n_ker) + w * (j + n_ker)] = v;
            norm = norm + v;
        }
    for (long q = 0; q < (long)w * w; ++q) ker[q] = ker[q] / norm;
}

int orl_add_source(orl_sim* s, int dir, double j_amp, double w0, double t0, double tau, double t_init, double t_final,
                   double phase, int n_ker, const int* ind, const double* ker) {
    lsource* q = &s->src[s->n_src++];
    const int w = 2 * n_ker + 1;
    const long np = s->dims == 1 ? w : (long)w * w;
    q->dir = dir; q->n_ker = n_ker; q->npts = (int)np;
    q->j_amp = j_amp; q->w0 = w0; q->t0 = t0; q->tau = tau; q->t_init = t_init; q->t_final = t_final; q->phase = phase;
    q->ind = (int*)malloc(sizeof(int) * s->dims * np); memcpy(q->ind, ind, sizeof(int) * s->dims * np);
    q->ker = (double*)malloc(sizeof(double) * np); memcpy(q->ker, ker, sizeof(double) * np);
    q->jmat = (double*)calloc((size_t)np, sizeof(double));
    return 0;
}

/* update_source (source_mod.f90:106-132) + source_interactions 1D / 2D
 * (interactions_mod.f90:32-100); J_mat one step past the kernel reads 0. */
void orl_inject_sources(orl_sim* s, double time) {
    const double c_src = s->dt_eps0 / s->dr / c0 / 2.0;
    for (int q = 0; q < s->n_src; ++q) {
        lsource* S = &s->src[q];
        const int n = S->n_ker, w = 2 * n + 1;
        const long np = S->npts;
        if (time >= S->t_init && time <= S->t_final) {
            const double x = (time - S->t0) / S->tau;
            const double envelope = exp(-(x * x));
            const double cos_t = cos(S->w0 * (time - S->t0) + S->phase);
            for (long p = 0; p < np; ++p) S->jmat[p] = S->j_amp * envelope * cos_t * S->ker[p];
        } else {
            for (long p = 0; p < np; ++p) S->jmat[p] = 0.0;
        }
        if (s->dims == 1) {
            for (int i = -n; i <= n; ++i) {
                const int ii = S->ind[i + n];
                AT(s->Ex, ii, 1, 1) = AT(s->Ex, ii, 1, 1) + c_src * S->jmat[i + n];
            }
            continue;
        }
        for (int i = -n; i <= n; ++i)
            for (int j = -n; j <= n; ++j) {
                const long p = (i + n) + w * (long)(j + n);
                const int ii = S->ind[p], jj = S->ind[p + np];
                if (S->dir == 1) {
                    const double jb = i + 1 <= n ? S->jmat[p + 1] : 0.0;
                    const double J_av = 0.5 * (S->jmat[p] + jb);
                    AT(s->Ex, ii, jj, 1) = AT(s->Ex, ii, jj, 1) + c_src * J_av;
                } else if (S->dir == 2) {
                    const double jb = j + 1 <= n ? S->jmat[p + w] : 0.0;
                    const double J_av = 0.5 * (S->jmat[p] + jb);
                    AT(s->Ey, ii, jj, 1) = AT(s->Ey, ii, jj, 1) + c_src * J_av;
                } else {
                    const double J_av = S->jmat[p];
                    AT(s->Ez, ii, jj, 1) = AT(s->Ez, ii, jj, 1) + c_src * J_av;
                }
            }
    }
}

/* detectors: write_1D_field / write_2D_field stencils (write_fields_subs_mod.f90:88-628) */
int orl_add_detector(orl_sim* s, int field, int nd, const int* idx) {
    const int q = s->n_det++;
    s->det_field[q] = field; s->det_nd[q] = nd;
    s->det_idx[q] = (int*)malloc(sizeof(int) * s->dims * (nd > 0 ? nd : 1));
    memcpy(s->det_idx[q], idx, sizeof(int) * s->dims * nd);
    s->n_samples += nd;
    return 0;
}

void orl_sample(const orl_sim* s, double* out) {
    long w = 0;
    for (int q = 0; q < s->n_det; ++q) {
        const int nd = s->det_nd[q];
        for (int n = 0; n < nd; ++n) {
            const int i = s->det_idx[q][n], j = s->dims == 1 ? 1 : s->det_idx[q][n + nd];
            double v = 0.0;
            if (s->dims == 1) {
                v = s->det_field[q] == 1 ? AT(s->Ex, i, 1, 1) : 0.5 * (AT(s->Hy, i - 1, 1, 1) + AT(s->Hy, i, 1, 1));
            } else {
                switch (s->det_field[q]) {
                    case 1: v = 0.5 * (AT(s->Ex, i - 1, j, 1) + AT(s->Ex, i, j, 1)); break;
                    case 2: v = 0.5 * (AT(s->Ey, i, j - 1, 1) + AT(s->Ey, i, j, 1)); break;
                    case 3: v = AT(s->Ez, i, j, 1); break;
                    case 4: v = 0.5 * (AT(s->Hx, i, j, 1) + AT(s->Hx, i, j - 1, 1)); break;
                    case 5: v = 0.5 * (AT(s->Hy, i, j, 1) + AT(s->Hy, i - 1, j, 1)); break;
                    case 6: v = 0.25 * (AT(s->Hz, i, j, 1) + AT(s->Hz, i - 1, j - 1, 1) + AT(s->Hz, i - 1, j, 1) + AT(s->Hz, i, j - 1, 1)); break;
                }
            }
            out[w++] = v;
        }
    }
}

long orl_n_samples(const orl_sim* s) { return s->n_samples; }

/* the driver loop's step order (mxim_mxll.f90:92-132) */
void orl_run(orl_sim* s, int tt_first, int n_steps, int record_every, double* rec, int* rows) {
    int r = 0;
    for (int tt = tt_first; tt < tt_first + n_steps; ++tt) {
        const double time = s->dt * ((double)tt - 0.5);
        orl_step_h(s);
        orl_inject_sources(s, time);
        orl_step_e(s);
        if (record_every > 0 && tt % record_every == 0 && s->n_samples > 0) {
            orl_sample(s, rec + (long)r * s->n_samples);
            ++r;
        }
    }
    if (rows) *rows = r;
}

/* copy one field with the reference's bounds, Fortran order (to_sim: host -> oracle) */
void orl_copy_field(orl_sim* s, int field, double* host, int to_sim) {
    arr3* a[6] = {&s->Ex, &s->Ey, &s->Ez, &s->Hx, &s->Hy, &s->Hz};
    arr3* f = a[field - 1];
    if (s->dims == 1 && field == 5) {   /* Hy(1:nz-1) */
        for (int k = 1; k <= s->nx - 1; ++k) {
            if (to_sim) AT(*f, k, 1, 1) = host[k - 1];
            else host[k - 1] = AT(*f, k, 1, 1);
        }
        return;
    }
    if (to_sim) memcpy(f->p, host, sizeof(double) * f->n);
    else memcpy(host, f->p, sizeof(double) * f->n);
}
