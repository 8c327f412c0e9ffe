// omx_lowdim.hpp -- the 2D (TMz / TEz / full) and 1D solvers of
// libomx_fdtd_hip (included at the end of omx_fdtd.hip, FP64).
//
// The 2D solver follows td_propagate_H_2Dfield / td_propagate_E_2Dfield
// (mxll_2D_mod.f90:450-864), the 1D one td_propagate_H_1Dfield /
// td_propagate_E_1Dfield (mxll_1D_mod.f90:249-352): one thread per cell,
// the bulk update, then the CPML psi of the faces the cell lies in (x, then
// y), in the reference's expression order; the classical media go through
// the same per-cell ADE update as the 3D engine (medium_update).  Fields are
// (n1+2) x (n2+2) arrays, index (i, j) -> i*(n2+2) + j (1D: n1+2 entries),
// which hold every reference array with its bounds (e.g. Ex(0:nx,1:ny+1));
// the CPML psi of a component lives at the cell it belongs to, in an array
// of the same shape.  Grids of these dimensions are small next to 3D ones,
// so the kernels are plain (no tiling); the 3D engine is the hot path.

namespace {

constexpr int kTMz = 1, kTEz = 2, kFull2D = 3;
// psi arrays: (component, face axis) of psix_Hyx, psiy_Hxy, psix_Ezx, psiy_Ezy
// (TMz) and psix_Hzx, psiy_Hzy, psix_Eyx, psiy_Exy (TEz) (mxll_2D_mod.f90:29-36)
enum { P_HYX, P_HXY, P_EZX, P_EZY, P_HZX, P_HZY, P_EYX, P_EXY, N_PSI };

struct LowArgs {
    double *ex, *ey, *ez, *hx, *hy, *hz;
    double* psi[N_PSI];
    const double *be, *ce, *bh, *ch;                  // 1-based CPML coefficients
    const double *dex, *dey, *dhx, *dhy;              // 1-based den arrays of the two axes
    const uint8_t *mx, *my, *mz;                      // effective medium per component and cell (null: none)
    const MedDev<double>* med;
    double *eold[3], *pd[3], *pl[3], *plold[3];       // dispersive state per component, dense
    long ps;                                          // pole stride (cells of the padded array)
    int n1, n2, s1;                                   // cells per axis, row stride (n2 + 2)
    int npml;
    unsigned cp;                                      // bit 0/1: x low/high face, 2/3: y low/high
    double dt_mu0, dt_eps0, dr;
};

// psi <- b psi + c rot, in place at o; returns the new psi
__device__ __forceinline__ double psi_upd(double* p, long o, double b, double c, double rot) {
    const double v = b * p[o] + c * rot;
    p[o] = v;
    return v;
}

__device__ __forceinline__ double med_e(const LowArgs& a, const uint8_t* mm, int c, long o, double e, double rot) {
    const int m = mm ? mm[o] : 0;
    return medium_update<double>(e, rot, a.dt_eps0 * rot, m, a.med, o, a.eold[c], a.pd[c], a.pl[c], a.plold[c], a.ps);
}

// H of the TMz family (Hx, Hy from Ez) and/or the TEz family (Hz from Ex, Ey).
__global__ void k2_h(LowArgs a, int mode) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)a.n1 * a.n2) return;
    const int i = (int)(t / a.n2) + 1, j = (int)(t % a.n2) + 1;
    const long o = (long)i * a.s1 + j, S = a.s1;
    const int np = a.npml, nx = a.n1, ny = a.n2;
    const bool xf = ((a.cp & 1u) && i <= np - 1) || ((a.cp & 2u) && i >= nx + 1 - np && i <= nx - 1);
    const bool yf = ((a.cp & 4u) && j <= np - 1) || ((a.cp & 8u) && j >= ny + 1 - np && j <= ny - 1);
    const int cx = i <= np - 1 ? i : nx - i, cy = j <= np - 1 ? j : ny - j;   // bh/ch index of the face cell
    if (mode & kTMz) {
        double hx = a.hx[o], hy = a.hy[o];
        hx = hx + a.dt_mu0 * ((a.ez[o] - a.ez[o + 1]) * a.dhy[j]);
        hy = hy + a.dt_mu0 * ((a.ez[o + S] - a.ez[o]) * a.dhx[i]);
        if (xf) hy = hy + a.dt_mu0 * psi_upd(a.psi[P_HYX], o, a.bh[cx], a.ch[cx], (a.ez[o + S] - a.ez[o]) / a.dr);
        if (yf) hx = hx + a.dt_mu0 * psi_upd(a.psi[P_HXY], o, a.bh[cy], a.ch[cy], (a.ez[o] - a.ez[o + 1]) / a.dr);
        a.hx[o] = hx; a.hy[o] = hy;
    }
    if (mode & kTEz) {
        double hz = a.hz[o];
        hz = hz + a.dt_mu0 * ((a.ey[o] - a.ey[o + S]) * a.dhx[i] + (a.ex[o + 1] - a.ex[o]) * a.dhy[j]);
        if (xf) hz = hz + a.dt_mu0 * psi_upd(a.psi[P_HZX], o, a.bh[cx], a.ch[cx], (a.ey[o] - a.ey[o + S]) / a.dr);
        if (yf) hz = hz + a.dt_mu0 * psi_upd(a.psi[P_HZY], o, a.bh[cy], a.ch[cy], (a.ex[o + 1] - a.ex[o]) / a.dr);
        a.hz[o] = hz;
    }
}

// E of the TMz family (Ez) and/or the TEz family (Ex, Ey): the medium
// update of each component, then the CPML psi of its faces.
__global__ void k2_e(LowArgs a, int mode) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)a.n1 * a.n2) return;
    const int i = (int)(t / a.n2) + 1, j = (int)(t % a.n2) + 1;
    const long o = (long)i * a.s1 + j, S = a.s1;
    const int np = a.npml, nx = a.n1, ny = a.n2;
    const bool xf = ((a.cp & 1u) && i >= 2 && i <= np) || ((a.cp & 2u) && i >= nx + 1 - np && i <= nx - 1);
    const bool yf = ((a.cp & 4u) && j >= 2 && j <= np) || ((a.cp & 8u) && j >= ny + 1 - np && j <= ny - 1);
    const int cx = i <= np ? i : nx + 1 - i, cy = j <= np ? j : ny + 1 - j;   // be/ce index of the face cell
    if (mode & kTMz) {
        const double rot = (a.hy[o] - a.hy[o - S]) * a.dex[i] + (a.hx[o - 1] - a.hx[o]) * a.dey[j];
        double ez = med_e(a, a.mz, 2, o, a.ez[o], rot);
        if (xf) ez = ez + a.dt_eps0 * psi_upd(a.psi[P_EZX], o, a.be[cx], a.ce[cx], (a.hy[o] - a.hy[o - S]) / a.dr);
        if (yf) ez = ez + a.dt_eps0 * psi_upd(a.psi[P_EZY], o, a.be[cy], a.ce[cy], (a.hx[o - 1] - a.hx[o]) / a.dr);
        a.ez[o] = ez;
    }
    if (mode & kTEz) {
        double ex = med_e(a, a.mx, 0, o, a.ex[o], (a.hz[o] - a.hz[o - 1]) * a.dey[j]);
        double ey = med_e(a, a.my, 1, o, a.ey[o], (a.hz[o - S] - a.hz[o]) * a.dex[i]);
        if (xf) ey = ey + a.dt_eps0 * psi_upd(a.psi[P_EYX], o, a.be[cx], a.ce[cx], (a.hz[o - S] - a.hz[o]) / a.dr);
        if (yf) ex = ex + a.dt_eps0 * psi_upd(a.psi[P_EXY], o, a.be[cy], a.ce[cy], (a.hz[o] - a.hz[o - 1]) / a.dr);
        a.ex[o] = ex; a.ey[o] = ey;
    }
}

}  // namespace

namespace {

// 1D (mxll_1D_mod.f90:249-352): Ex(1..nz) at k, Hy(1..nz-1) at k+1/2; Ex(1)
// and Ex(nz) are never updated.  The expressions keep the reference's
// association: Hy + (dt_mu0*(Ex(k)-Ex(k+1)))*den_hz(k), psi = b*psi +
// (c*(Ex(k)-Ex(k+1)))/dr.
__global__ void k1_h(LowArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int nz = a.n1, np = a.npml;
    if (k > nz - 1) return;
    double* ex = a.ex;
    double hy = a.hy[k] + a.dt_mu0 * (ex[k] - ex[k + 1]) * a.dhx[k];
    const bool lo = (a.cp & 1u) && k <= np - 1, hi = (a.cp & 2u) && k >= nz + 1 - np;
    if (lo || hi) {
        const int c = lo ? k : nz - k;
        double* p = a.psi[P_HYX];
        p[k] = a.bh[c] * p[k] + a.ch[c] * (ex[k] - ex[k + 1]) / a.dr;
        hy = hy + a.dt_mu0 * p[k];
    }
    a.hy[k] = hy;
}

__global__ void k1_e(LowArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int nz = a.n1, np = a.npml;
    if (k > nz - 1) return;
    const double rot = (a.hy[k - 1] - a.hy[k]) * a.dex[k];
    double ex = med_e(a, a.mx, 0, k, a.ex[k], rot);
    const bool lo = (a.cp & 1u) && k <= np, hi = (a.cp & 2u) && k >= nz + 1 - np;
    if (lo || hi) {
        const int c = lo ? k : nz + 1 - k;
        double* p = a.psi[P_EZX];
        p[k] = a.be[c] * p[k] + a.ce[c] * ((a.hy[k - 1] - a.hy[k]) / a.dr);
        ex = ex + a.dt_eps0 * p[k];
    }
    a.ex[k] = ex;
}

// Periodic copies of the non-MPI branches: dst <- src for the whole line
// (or column) of the padded array.  axis 0: dst/src are i indices, every j;
// axis 1: j indices, every i.
__global__ void k2_copy(double* f, int s1, int n_other, int axis, int dst, int src) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n_other) return;
    if (axis == 0) f[(long)dst * s1 + u] = f[(long)src * s1 + u];
    else f[(long)u * s1 + dst] = f[(long)u * s1 + src];
}

// Detector samples (write_2D_field / write_1D_field stencils,
// write_fields_subs_mod.f90:16-178,264-613): code = field (1..6) per
// sample, off = its cell.
__global__ void k2_sample(LowArgs a, int dims, const int* __restrict__ code, const long* __restrict__ off, int n,
                          double* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long o = off[q], S = a.s1;
    double v = 0.0;
    if (dims == 1) {
        v = code[q] == 1 ? a.ex[o] : 0.5 * (a.hy[o - 1] + a.hy[o]);
    } else {
        switch (code[q]) {
            case 1: v = 0.5 * (a.ex[o - S] + a.ex[o]); break;
            case 2: v = 0.5 * (a.ey[o - 1] + a.ey[o]); break;
            case 3: v = a.ez[o]; break;
            case 4: v = 0.5 * (a.hx[o] + a.hx[o - 1]); break;
            case 5: v = 0.5 * (a.hy[o] + a.hy[o - S]); break;
            default: v = 0.25 * (a.hz[o] + a.hz[o - S - 1] + a.hz[o - S] + a.hz[o - 1]); break;
        }
    }
    out[q] = v;
}

}  // namespace

struct omxl_ctx {
    omxl_desc d;
    int device = 0;
    hipStream_t s = nullptr;
    int s1 = 1;                   // row stride (n2 + 2)
    long ncell = 0;               // (n1+2) * (n2+2)
    unsigned cp = 0;              // CPML face bits
    double* f[6] = {};            // Ex Ey Ez Hx Hy Hz (the mode's components)
    double* psi[N_PSI] = {};
    double* coef = nullptr;       // be ce bh ch (npml+1 each), den_x e/h, den_y e/h (n+2 each), spare
    long co_be = 0, co_ce = 0, co_bh = 0, co_ch = 0, co_dex = 0, co_dhx = 0, co_dey = 0, co_dhy = 0;
    // media
    int n_media = 0, maxp = 0;
    omx_medium* d_med_in = nullptr;
    omx_medium_coeffs* d_med_co = nullptr;
    MedDev<double>* d_med = nullptr;
    uint8_t* d_m[3] = {};         // effective medium per component
    double* st[3][4] = {};        // eold pd pl plold per component
    // sources, detectors
    struct Src { int dir; double j_amp, w0, t0, tau, t_init, t_final, phase; int npts; long* d_off; double *d_ka, *d_kb; int field; };
    std::vector<Src> src;
    std::vector<int> h_code;
    std::vector<long> h_off;
    int* d_code = nullptr;
    long* d_off = nullptr;
    double* d_samp = nullptr;
    double* d_rec = nullptr;      // omxl_run's rows of samples
    long rec_cap = 0;             // doubles held by d_rec
    int n_det_dev = 0;            // samples the device lists hold
    bool det_dirty = false;
    long device_bytes = 0;
};

namespace {

int lalloc(omxl_ctx* c, void** p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(OMX_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    c->device_bytes += (long)bytes;
    HIPCHK(hipMemsetAsync(*p, 0, bytes, c->s));
    return OMX_OK;
}

bool low_has(const omxl_ctx* c, int field) {   // 1-based field code
    if (c->d.dims == 1) return field == OMX_EX || field == OMX_HY;
    const bool tm = c->d.mode & kTMz, te = c->d.mode & kTEz;
    return (field == OMX_EZ || field == OMX_HX || field == OMX_HY) ? tm : te;
}

LowArgs low_args(omxl_ctx* c) {
    LowArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ex = c->f[0]; a.ey = c->f[1]; a.ez = c->f[2]; a.hx = c->f[3]; a.hy = c->f[4]; a.hz = c->f[5];
    for (int q = 0; q < N_PSI; ++q) a.psi[q] = c->psi[q];
    const double* k = c->coef;
    a.be = k + c->co_be; a.ce = k + c->co_ce; a.bh = k + c->co_bh; a.ch = k + c->co_ch;
    a.dex = k + c->co_dex; a.dhx = k + c->co_dhx; a.dey = k + c->co_dey; a.dhy = k + c->co_dhy;
    a.mx = c->d_m[0]; a.my = c->d_m[1]; a.mz = c->d_m[2];
    if (c->d.dims == 1) a.mx = c->d_m[0];
    a.med = c->d_med;
    for (int q = 0; q < 3; ++q) { a.eold[q] = c->st[q][0]; a.pd[q] = c->st[q][1]; a.pl[q] = c->st[q][2]; a.plold[q] = c->st[q][3]; }
    a.ps = c->ncell;
    a.n1 = c->d.n1; a.n2 = c->d.dims == 1 ? 1 : c->d.n2; a.s1 = c->s1;
    a.npml = c->d.npml;
    a.cp = c->cp;
    a.dt_mu0 = c->d.dt / kMu0; a.dt_eps0 = c->d.dt / kEps0; a.dr = c->d.dr;
    return a;
}

// Reference bounds of a field in the padded array (lo/hi per axis, 2D;
// 1D: axis 0 only).
void low_bounds(const omxl_ctx* c, int field, int lo[2], int hi[2]) {
    const int n1 = c->d.n1, n2 = c->d.n2;
    if (c->d.dims == 1) {
        lo[0] = 1; hi[0] = field == OMX_EX ? n1 : n1 - 1;
        lo[1] = hi[1] = 0;
        return;
    }
    // Ex(0:nx,1:ny+1) Ey(1:nx+1,0:ny) Ez(1:nx+1,1:ny+1) Hx(1:nx,0:ny) Hy(0:nx,1:ny) Hz(0:nx,0:ny)
    static const int tbl[6][4] = {{0, 0, 1, 1}, {1, 1, 0, 0}, {1, 1, 1, 1}, {1, 0, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 0}};
    const int* t = tbl[field - 1];
    lo[0] = t[0]; hi[0] = n1 + t[1];
    lo[1] = t[2]; hi[1] = n2 + t[3];
}

}  // namespace

int omxl_create(const omxl_desc* desc, omxl_ctx** out) {
    if (!desc || !out) return fail(OMX_ERR_ARG, "null argument");
    *out = nullptr;
    const omxl_desc& d = *desc;
    if (d.dims != 1 && d.dims != 2) return fail(OMX_ERR_ARG, "dims must be 1 or 2");
    if (d.precision != OMX_FP64) return fail(OMX_ERR_ARG, "the 1D and 2D solvers run in FP64 (precision 8)");
    if (d.n1 < 3 || (d.dims == 2 && d.n2 < 3)) return fail(OMX_ERR_ARG, "grid must be at least 3 cells per axis");
    if (d.dims == 2 && (d.mode < kTMz || d.mode > kFull2D)) return fail(OMX_ERR_ARG, "2D mode must be 1 (TMz), 2 (TEz) or 3 (full)");
    if (!(d.dt > 0.0) || !(d.dr > 0.0)) return fail(OMX_ERR_ARG, "dt and dr must be positive");
    unsigned cp = 0;
    for (int a = 0; a < d.dims; ++a) {
        const int b = d.boundaries[a];
        if (b != OMX_BC_CLOSE && b != OMX_BC_PERIODIC && b != OMX_BC_CPML)
            return fail(OMX_ERR_ARG, "boundary code of axis " + std::to_string(a + 1) + " must be 1, 2 or 3");
        if (b == OMX_BC_CPML) {
            if (d.npml < 2) return fail(OMX_ERR_ARG, "CPML needs npml >= 2");
            if (2 * d.npml > (a == 0 ? d.n1 : d.n2)) return fail(OMX_ERR_ARG, "CPML thicker than half the grid");
            cp |= 3u << (2 * a);
        }
    }
    omxl_ctx* c = new omxl_ctx();
    c->d = d;
    if (d.dims == 1) { c->d.n2 = 1; c->d.mode = 0; }
    c->cp = cp;
    auto bail = [&](int r) { omxl_destroy(c); return r; };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(OMX_ERR_HIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(OMX_ERR_HIP, "hipStreamCreate failed"));
    c->s1 = d.dims == 1 ? 1 : d.n2 + 2;
    c->ncell = d.dims == 1 ? (long)d.n1 + 2 : (long)(d.n1 + 2) * (d.n2 + 2);
    int rc = OMX_OK;
    for (int q = 1; q <= 6 && !rc; ++q)
        if (low_has(c, q)) rc = lalloc(c, (void**)&c->f[q - 1], sizeof(double) * c->ncell);
    if (rc) return bail(rc);
    if (cp) {
        int need[N_PSI] = {0};
        if (d.dims == 1) { need[P_HYX] = need[P_EZX] = 1; }
        else {
            if (d.mode & kTMz) need[P_HYX] = need[P_HXY] = need[P_EZX] = need[P_EZY] = 1;
            if (d.mode & kTEz) need[P_HZX] = need[P_HZY] = need[P_EYX] = need[P_EXY] = 1;
        }
        for (int q = 0; q < N_PSI && !rc; ++q)
            if (need[q]) rc = lalloc(c, (void**)&c->psi[q], sizeof(double) * c->ncell);
        if (rc) return bail(rc);
    }
    // coefficients: CPML tables and den arrays of the two axes, computed as
    // for the 3D grid (the reference's 1D and 2D init compute the same
    // profiles, mxll_2D_mod.f90:262-378, mxll_1D_mod.f90:138-210)
    const int np = std::max(d.npml, 2), n1 = d.n1, n2 = d.dims == 1 ? 1 : d.n2;
    c->co_be = 0; c->co_ce = np + 1; c->co_bh = 2 * (np + 1); c->co_ch = c->co_bh + np + 1;
    c->co_dex = c->co_ch + np + 1; c->co_dhx = c->co_dex + n1 + 2;
    c->co_dey = c->co_dhx + n1 + 2; c->co_dhy = c->co_dey + n2 + 2;
    const long ncoef = c->co_dhy + n2 + 2 + 4;   // + a spare den pair for the unused third axis
    rc = lalloc(c, (void**)&c->coef, sizeof(double) * ncoef);
    if (rc) return bail(rc);
    CpmlProfile pr;
    pr.npml = d.npml; pr.dt = d.dt; pr.dr = d.dr;
    const double sig_to_au = ((kCToAu * kCToAu) * kSecToAu) / (kKgToAu * ((kMToAu * kMToAu) * kMToAu));
    pr.sig_max = (double)(0.8f * (float)(kGradeM + 1)) / (d.dr * std::sqrt(kMu0 / kEps0));
    pr.alpha_scale = kAlphaCPML * sig_to_au;
    double* k = c->coef;
    const int kmax = std::max(std::max(n1, n2), np);
    hipLaunchKernelGGL(k_cpml_coeffs, dim3(blocks_for(kmax, 128)), dim3(128), 0, c->s, pr, n1, n2, 1, cp,
                       k + c->co_be, k + c->co_ce, k + c->co_bh, k + c->co_ch, k + c->co_dex, k + c->co_dey,
                       k + ncoef - 4, k + c->co_dhx, k + c->co_dhy, k + ncoef - 2);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->s) != hipSuccess)
        return bail(fail(OMX_ERR_HIP, "coefficient set-up failed"));
    *out = c;
    return OMX_OK;
}

int omxl_destroy(omxl_ctx* c) {
    if (!c) return OMX_OK;
    if (c->s) (void)hipStreamSynchronize(c->s);
    for (auto& p : c->f) dfreet(p);
    for (auto& p : c->psi) dfreet(p);
    dfreet(c->coef);
    dfreet(c->d_med_in); dfreet(c->d_med_co); dfreet(c->d_med);
    for (auto& p : c->d_m) dfreet(p);
    for (auto& r : c->st) for (auto& p : r) dfreet(p);
    for (auto& s : c->src) { dfreet(s.d_off); dfreet(s.d_ka); dfreet(s.d_kb); }
    dfreet(c->d_code); dfreet(c->d_off); dfreet(c->d_samp); dfreet(c->d_rec);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return OMX_OK;
}

int omxl_set_media(omxl_ctx* c, int32_t n_media, const omx_medium* table, const int8_t* aux) {
    CHECK_CTX(c);
    if (c->n_media) return fail(OMX_ERR_STATE, "media already set for this context");
    if (n_media < 1 || n_media > OMX_MAX_MEDIA) return fail(OMX_ERR_RANGE, "n_media must be in 1..15");
    if (!table || !aux) return fail(OMX_ERR_ARG, "null media table or aux map");
    std::vector<omx_medium> tab(n_media + 1);
    std::memset(tab.data(), 0, sizeof(omx_medium) * tab.size());
    for (int m = 1; m <= n_media; ++m) {
        tab[m] = table[m - 1];
        if (tab[m].kind < OMX_MEDIUM_DIELECTRIC || tab[m].kind > OMX_MEDIUM_DL)
            return fail(OMX_ERR_ARG, "medium " + std::to_string(m) + ": unknown kind");
        if (tab[m].kind == OMX_MEDIUM_DL && (tab[m].n_poles < 1 || tab[m].n_poles > OMX_MAX_POLES))
            return fail(OMX_ERR_RANGE, "medium " + std::to_string(m) + ": n_poles must be in 1..8");
        if (!(tab[m].eps_r > 0.0)) return fail(OMX_ERR_ARG, "medium " + std::to_string(m) + ": eps_r must be > 0");
        if (tab[m].kind == OMX_MEDIUM_DL) c->maxp = std::max(c->maxp, (int)tab[m].n_poles);
    }
    const int n1 = c->d.n1, n2 = c->d.n2;
    const long ncell = (long)n1 * n2;
    for (long q = 0; q < ncell; ++q)
        if (aux[q] < 0 || aux[q] > n_media) return fail(OMX_ERR_ARG, "aux map value out of 0..n_media");
    auto disp = [&](int m) { return m && (tab[m].kind == OMX_MEDIUM_DRUDE || tab[m].kind == OMX_MEDIUM_DL); };
    // effective medium per component (the padded layout): the medium of the
    // component's staggered map when it is a Drude / Drude-Lorentz one, else
    // the dielectric that sets the component's eps (or none)
    std::vector<uint8_t> eff[3];
    for (auto& e : eff) e.assign(c->ncell, 0);
    if (c->d.dims == 1) {
        // read_init_media_1D (classical_medium_mod.f90:203-337): the cell's medium
        for (int i = 1; i <= n1; ++i) eff[0][i] = (uint8_t)aux[i - 1];
    } else {
        // read_init_media_2D (classical_medium_mod.f90:484-538): staggered
        // map per component, and eps of every component from component 1's
        std::vector<int> map(3 * (size_t)ncell, 0);
        auto A = [&](int i, int j) { return (i < 1 || i > n1 || j < 1 || j > n2) ? 0 : (int)aux[(i - 1) + (long)n1 * (j - 1)]; };
        auto M = [&](int i, int j, int q) -> int& { return map[(size_t)q * ncell + (i - 1) + (size_t)n1 * (j - 1)]; };
        for (int n = 1; n <= n_media; ++n)
            for (int j = 1; j <= n2; ++j)
                for (int i = 1; i <= n1; ++i) {
                    const int a0 = A(i, j);
                    if (i == n1 && a0 == n) { M(i, j, 0) = n; }
                    else if (j == n2 && a0 == n) { M(i, j, 0) = n; }
                    else {
                        if (a0 == n && A(i + 1, j) == n) M(i, j, 0) = n;
                        else if (a0 != 0 && A(i + 1, j) == n) M(i, j, 0) = n;
                        if (a0 == n && A(i, j + 1) == n) M(i, j, 1) = n;
                        else if (a0 != 0 && A(i, j + 1) == n) M(i, j, 1) = n;
                        if (a0 == n) M(i, j, 2) = n;
                    }
                }
        for (int j = 1; j <= n2; ++j)
            for (int i = 1; i <= n1; ++i) {
                const int m1 = M(i, j, 0);
                const int diel = (m1 && tab[m1].kind == OMX_MEDIUM_DIELECTRIC) ? m1 : 0;
                for (int q = 0; q < 3; ++q) {
                    const int mq = M(i, j, q);
                    eff[q][(long)i * c->s1 + j] = (uint8_t)(disp(mq) ? mq : diel);
                }
            }
    }
    int rc = lalloc(c, (void**)&c->d_med_in, sizeof(omx_medium) * (n_media + 1));
    if (!rc) rc = lalloc(c, (void**)&c->d_med_co, sizeof(omx_medium_coeffs) * (n_media + 1));
    if (!rc) rc = lalloc(c, (void**)&c->d_med, sizeof(MedDev<double>) * (n_media + 1));
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->d_med_in, tab.data(), sizeof(omx_medium) * tab.size(), hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_medium_coeffs<double>, dim3(1), dim3(64), 0, c->s, c->d_med_in, n_media, c->d.dt,
                       c->d.dt / kEps0, c->d_med_co, c->d_med);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 3; ++q) {
        const int field = q + 1;
        if (!low_has(c, field)) continue;
        rc = lalloc(c, (void**)&c->d_m[q], c->ncell);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(c->d_m[q], eff[q].data(), c->ncell, hipMemcpyHostToDevice, c->s));
        const int p = std::max(c->maxp, 1);
        const size_t sz[4] = {1, 1, (size_t)p, (size_t)p};
        for (int w = 0; w < 4 && !rc; ++w) rc = lalloc(c, (void**)&c->st[q][w], sizeof(double) * sz[w] * c->ncell);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->s));
    c->n_media = n_media;
    return OMX_OK;
}

int omxl_step_h(omxl_ctx* c) {
    CHECK_CTX(c);
    const LowArgs a = low_args(c);
    const int n1 = c->d.n1, n2 = c->d.n2;
    if (c->d.dims == 1) {
        if (c->d.boundaries[0] == OMX_BC_PERIODIC)   // Ex(nz+1) = Ex(1) (mxll_1D_mod.f90:262-264)
            hipLaunchKernelGGL(k2_copy, dim3(1), dim3(64), 0, c->s, c->f[0], 1, 1, 0, n1 + 1, 1);
        hipLaunchKernelGGL(k1_h, dim3(blocks_for(n1, 256)), dim3(256), 0, c->s, a);
    } else {
        const int s1 = c->s1;
        const bool px = c->d.boundaries[0] == OMX_BC_PERIODIC, py = c->d.boundaries[1] == OMX_BC_PERIODIC;
        // mxll_2D_mod.f90:475-481 (TMz) and :561-567 (TEz)
        if (c->d.mode & kTMz) {
            if (px) hipLaunchKernelGGL(k2_copy, dim3(blocks_for(n2 + 2, 64)), dim3(64), 0, c->s, c->f[2], s1, n2 + 2, 0, n1 + 1, 1);
            if (py) hipLaunchKernelGGL(k2_copy, dim3(blocks_for(n1 + 2, 64)), dim3(64), 0, c->s, c->f[2], s1, n1 + 2, 1, n2 + 1, 1);
        }
        if (c->d.mode & kTEz) {
            if (px) hipLaunchKernelGGL(k2_copy, dim3(blocks_for(n2 + 2, 64)), dim3(64), 0, c->s, c->f[1], s1, n2 + 2, 0, n1 + 1, 1);
            if (py) hipLaunchKernelGGL(k2_copy, dim3(blocks_for(n1 + 2, 64)), dim3(64), 0, c->s, c->f[0], s1, n1 + 2, 1, n2 + 1, 1);
        }
        hipLaunchKernelGGL(k2_h, dim3(blocks_for((long)n1 * n2, 256)), dim3(256), 0, c->s, a, c->d.mode);
    }
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

int omxl_step_e(omxl_ctx* c) {
    CHECK_CTX(c);
    const LowArgs a = low_args(c);
    const int n1 = c->d.n1, n2 = c->d.n2;
    if (c->d.dims == 1) {
        hipLaunchKernelGGL(k1_e, dim3(blocks_for(n1, 256)), dim3(256), 0, c->s, a);
    } else {
        const int s1 = c->s1;
        const bool px = c->d.boundaries[0] == OMX_BC_PERIODIC, py = c->d.boundaries[1] == OMX_BC_PERIODIC;
        // mxll_2D_mod.f90:666-672 (TMz) and :757-763 (TEz)
        if (c->d.mode & kTMz) {
            if (px) hipLaunchKernelGGL(k2_copy, dim3(blocks_for(n2 + 2, 64)), dim3(64), 0, c->s, c->f[4], s1, n2 + 2, 0, 0, n1);
            if (py) hipLaunchKernelGGL(k2_copy, dim3(blocks_for(n1 + 2, 64)), dim3(64), 0, c->s, c->f[3], s1, n1 + 2, 1, 0, n2);
        }
        if (c->d.mode & kTEz) {
            if (px) hipLaunchKernelGGL(k2_copy, dim3(blocks_for(n2 + 2, 64)), dim3(64), 0, c->s, c->f[5], s1, n2 + 2, 0, 0, n1);
            if (py) hipLaunchKernelGGL(k2_copy, dim3(blocks_for(n1 + 2, 64)), dim3(64), 0, c->s, c->f[5], s1, n1 + 2, 1, 0, n2);
        }
        hipLaunchKernelGGL(k2_e, dim3(blocks_for((long)n1 * n2, 256)), dim3(256), 0, c->s, a, c->d.mode);
    }
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

int omxl_add_source(omxl_ctx* c, int32_t dir, double j_amp, double w0, double t0, double tau, double t_init,
                    double t_final, double phase, int32_t n_ker, const int32_t* ind, const double* ker) {
    CHECK_CTX(c);
    const int dims = c->d.dims;
    if (n_ker < 0 || !ind || !ker) return fail(OMX_ERR_ARG, "bad source kernel");
    if (dir < 1 || dir > 3 || !low_has(c, dir) || (dims == 1 && dir != 1))
        return fail(OMX_ERR_ARG, "source direction has no field in this grid");
    const int w = 2 * n_ker + 1;
    const long np = dims == 1 ? w : (long)w * w;
    std::vector<long> off(np);
    std::vector<double> ka(np), kb(np);
    for (long p = 0; p < np; ++p) {
        const int i = (int)(p % w), j = (int)(p / w);
        const int ii = ind[p], jj = dims == 1 ? 0 : ind[p + np];
        if (ii < 1 || ii > c->d.n1 || (dims == 2 && (jj < 1 || jj > c->d.n2)))
            return fail(OMX_ERR_RANGE, "source point outside the grid");
        off[p] = dims == 1 ? ii : (long)ii * c->s1 + jj;
        // 1D Ex and 2D Ez take J_mat of the point itself (interactions_mod.f90:43-50,89-100);
        // 2D Ex / Ey the mean with the next point along x / y (:60-88; zero past
        // the kernel's edge).  k_source adds 1/2 (J ka + J kb): 2 ker and 0
        // give J ker exactly.
        if (dims == 1 || dir == 3) {
            ka[p] = 2.0 * ker[p];
            kb[p] = 0.0;
        } else {
            ka[p] = ker[p];
            const int ni = dir == 1 ? i + 1 : i, nj = dir == 2 ? j + 1 : j;
            kb[p] = (ni < w && nj < w) ? ker[ni + (long)w * nj] : 0.0;
        }
    }
    omxl_ctx::Src s;
    s.dir = dir; s.j_amp = j_amp; s.w0 = w0; s.t0 = t0; s.tau = tau; s.t_init = t_init; s.t_final = t_final;
    s.phase = phase; s.npts = (int)np; s.field = dir;
    s.d_off = nullptr; s.d_ka = nullptr; s.d_kb = nullptr;
    int rc = lalloc(c, (void**)&s.d_off, sizeof(long) * np);
    if (!rc) rc = lalloc(c, (void**)&s.d_ka, sizeof(double) * np);
    if (!rc) rc = lalloc(c, (void**)&s.d_kb, sizeof(double) * np);
    if (rc) { dfreet(s.d_off); dfreet(s.d_ka); dfreet(s.d_kb); return rc; }
    HIPCHK(hipMemcpyAsync(s.d_off, off.data(), sizeof(long) * np, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(s.d_ka, ka.data(), sizeof(double) * np, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(s.d_kb, kb.data(), sizeof(double) * np, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->src.push_back(s);
    return OMX_OK;
}

int omxl_inject_sources(omxl_ctx* c, double time) {
    CHECK_CTX(c);
    for (auto& s : c->src) {
        if (!(time >= s.t_init && time <= s.t_final)) continue;
        SourceParams sp;
        sp.j_amp = s.j_amp; sp.w0 = s.w0; sp.t0 = s.t0; sp.tau = s.tau; sp.t_init = s.t_init; sp.t_final = s.t_final;
        sp.phase = s.phase; sp.dt = c->d.dt;
        sp.c_src = ((c->d.dt / kEps0) / c->d.dr / kC0) / 2.0;   // dt_eps0/dr/c0/2 (interactions_mod.f90:32)
        hipLaunchKernelGGL(k_source<double>, dim3(blocks_for(s.npts, 256)), dim3(256), 0, c->s, c->f[s.field - 1],
                           s.d_off, s.d_ka, s.d_kb, s.npts, sp, time, (double* const*)nullptr);
    }
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

int omxl_add_detector(omxl_ctx* c, int32_t field, int32_t nd, const int32_t* idx) {
    CHECK_CTX(c);
    if (field < OMX_EX || field > OMX_HZ || !low_has(c, field)) return fail(OMX_ERR_ARG, "field not present in this grid");
    if (nd < 0 || (nd > 0 && !idx)) return fail(OMX_ERR_ARG, "bad detector index list");
    for (int q = 0; q < nd; ++q) {
        const int ii = idx[q], jj = c->d.dims == 1 ? 0 : idx[q + nd];
        if (ii < 1 || ii > c->d.n1 || (c->d.dims == 2 && (jj < 1 || jj > c->d.n2)))
            return fail(OMX_ERR_RANGE, "detector index outside the grid");
        c->h_code.push_back(field);
        c->h_off.push_back(c->d.dims == 1 ? ii : (long)ii * c->s1 + jj);
    }
    c->det_dirty = true;
    return OMX_OK;
}

namespace {

int upload_detectors(omxl_ctx* c) {
    const int n = (int)c->h_code.size();
    if (c->det_dirty) {
        HIPCHK(hipStreamSynchronize(c->s));
        if (c->d_code) c->device_bytes -= (long)(sizeof(int) + sizeof(long) + sizeof(double)) * c->n_det_dev;
        dfreet(c->d_code); dfreet(c->d_off); dfreet(c->d_samp);
        c->n_det_dev = n;
        int rc = lalloc(c, (void**)&c->d_code, sizeof(int) * n);
        if (!rc) rc = lalloc(c, (void**)&c->d_off, sizeof(long) * n);
        if (!rc) rc = lalloc(c, (void**)&c->d_samp, sizeof(double) * n);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(c->d_code, c->h_code.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->s));
        HIPCHK(hipMemcpyAsync(c->d_off, c->h_off.data(), sizeof(long) * n, hipMemcpyHostToDevice, c->s));
        c->det_dirty = false;
    }
    return OMX_OK;
}

}  // namespace

int omxl_sample(omxl_ctx* c, double* out) {
    CHECK_CTX(c);
    const int n = (int)c->h_code.size();
    if (n == 0) return OMX_OK;
    if (!out) return fail(OMX_ERR_ARG, "null output buffer");
    int rc = upload_detectors(c);
    if (rc) return rc;
    hipLaunchKernelGGL(k2_sample, dim3(blocks_for(n, 256)), dim3(256), 0, c->s, low_args(c), c->d.dims, c->d_code,
                       c->d_off, n, c->d_samp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, c->d_samp, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return OMX_OK;
}

int omxl_copy_field(omxl_ctx* c, int32_t field, double* host, int32_t to_device) {
    CHECK_CTX(c);
    if (field < 1 || field > 6 || !host || !low_has(c, field)) return fail(OMX_ERR_ARG, "bad field or null buffer");
    int lo[2], hi[2];
    low_bounds(c, field, lo, hi);
    std::vector<double> stage(c->ncell);
    HIPCHK(hipMemcpyAsync(stage.data(), c->f[field - 1], sizeof(double) * c->ncell, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    const long e0 = hi[0] - lo[0] + 1;
    for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
            const long h = (i - lo[0]) + e0 * (j - lo[1]);
            const long o = (long)i * c->s1 + j;
            if (to_device) stage[o] = host[h];
            else host[h] = stage[o];
        }
    if (to_device) {
        HIPCHK(hipMemcpyAsync(c->f[field - 1], stage.data(), sizeof(double) * c->ncell, hipMemcpyHostToDevice, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
    }
    return OMX_OK;
}

int omxl_synchronize(omxl_ctx* c) {
    CHECK_CTX(c);
    HIPCHK(hipStreamSynchronize(c->s));
    return OMX_OK;
}

int omxl_run(omxl_ctx* c, int32_t tt_first, int32_t n_steps, int32_t record_every, double* out) {
    CHECK_CTX(c);
    if (n_steps < 0) return fail(OMX_ERR_ARG, "n_steps must be >= 0");
    const int n = (int)c->h_code.size();
    long rows = 0;
    if (record_every > 0 && n > 0)
        for (long tt = tt_first; tt < (long)tt_first + n_steps; ++tt) rows += tt % record_every == 0;
    if (rows > 0 && !out) return fail(OMX_ERR_ARG, "null record buffer");
    if (rows > 0) {
        int rc = upload_detectors(c);
        if (rc) return rc;
        if (rows * n > c->rec_cap) {
            HIPCHK(hipStreamSynchronize(c->s));
            c->device_bytes -= (long)sizeof(double) * c->rec_cap;
            dfreet(c->d_rec);
            c->rec_cap = 0;
            rc = lalloc(c, (void**)&c->d_rec, sizeof(double) * rows * n);
            if (rc) return rc;
            c->rec_cap = rows * n;
        }
    }
    // the step order of the driver loop (mxim_mxll.f90:92-132): H, the
    // sources at t = dt (tt - 1/2), E, then the detectors
    long r = 0;
    for (long tt = tt_first; tt < (long)tt_first + n_steps; ++tt) {
        int rc = omxl_step_h(c);
        if (!rc) rc = omxl_inject_sources(c, c->d.dt * ((double)tt - 0.5));
        if (!rc) rc = omxl_step_e(c);
        if (rc) return rc;
        if (rows > 0 && tt % record_every == 0) {
            hipLaunchKernelGGL(k2_sample, dim3(blocks_for(n, 256)), dim3(256), 0, c->s, low_args(c), c->d.dims,
                               c->d_code, c->d_off, n, c->d_rec + r * n);
            ++r;
        }
    }
    HIPCHK(hipGetLastError());
    if (rows > 0)
        HIPCHK(hipMemcpyAsync(out, c->d_rec, sizeof(double) * rows * n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return OMX_OK;
}

int omxl_device_bytes(omxl_ctx* c, int64_t* bytes) {
    CHECK_CTX(c);
    if (bytes) *bytes = c->device_bytes;
    return OMX_OK;
}
