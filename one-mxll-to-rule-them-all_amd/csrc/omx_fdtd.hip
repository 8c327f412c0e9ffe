// omx_fdtd.hip -- C-ABI of libomx_fdtd_hip (declared in include/omx_fdtd.h).
//
// Host side of the MI355X FDTD engine: context lifecycle, coefficient and
// media set-up, the per-step launch sequence of the reference driver loop
// (src/FDTD/mxim_mxll.f90:92-132) and the device-resident couplings
// (sources, detectors, embedded quantum systems, TF/SF, energy).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "omx_kernels.hpp"

using namespace omx;

// Workgroups of the energy reduction (k_energy): each leaves three partial
// sums (FP64, compensated FP32 and its compensation) for k_energy_final.
constexpr int kEnergyBlocks = 1024;

// Scratch of omx3d_energy_noalloc: the partial sums of the energy reduction,
// 3 * kEnergyBlocks doubles (24 KiB) on every GPU the library runs on, held
// until the process ends.  Host code never uses this name as an address: the
// device address of each GPU's copy comes from hipGetSymbolAddress (below).
static __device__ __align__(256) double scratch[3 * kEnergyBlocks];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return fail(OMX_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHECK_CTX(c) \
    do {             \
        if (!(c)) return fail(OMX_ERR_ARG, "null context"); \
    } while (0)

int dalloc(omx3d_ctx* c, void** p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(OMX_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    c->device_bytes += (long)bytes;
    e = hipMemsetAsync(*p, 0, bytes, c->s);
    if (e != hipSuccess) return fail(OMX_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    return OMX_OK;
}

void dfree(void*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}
template <typename P>
void dfreet(P*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

inline unsigned blocks_for(long n, int b) { return (unsigned)((n + b - 1) / b); }

// Bytes of the compact Drude / DL state (YeeArgs::dstate).
size_t disp_state_bytes(const omx3d_ctx* c) {
    return c->esz * 3 * (size_t)(2 + 2 * c->maxp) * (size_t)std::max(c->n_disp, 1);
}

// Kernel argument block for precision T.
template <typename T>
YeeArgs<T> make_args(omx3d_ctx* c) {
    YeeArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    a.ex = (T*)c->f[0]; a.ey = (T*)c->f[1]; a.ez = (T*)c->f[2];
    a.hx = (T*)c->f[3]; a.hy = (T*)c->f[4]; a.hz = (T*)c->f[5];
    const T* tc = (const T*)c->tcoef;
    a.be = tc + c->co_be; a.ced = tc + c->co_ce; a.bh = tc + c->co_bh; a.chd = tc + c->co_ch;
    a.den_ex = tc + c->co_den[0]; a.den_ey = tc + c->co_den[1]; a.den_ez = tc + c->co_den[2];
    a.den_hx = tc + c->co_den[3]; a.den_hy = tc + c->co_den[4]; a.den_hz = tc + c->co_den[5];
    a.cdex = tc + c->co_cd[0]; a.cdey = tc + c->co_cd[1]; a.cdez = tc + c->co_cd[2];
    a.cdhx = tc + c->co_cd[3]; a.cdhy = tc + c->co_cd[4]; a.cdhz = tc + c->co_cd[5];
    for (int q = 0; q < 6; ++q) { a.ph[q] = (T*)c->psi_h[q]; a.pe[q] = (T*)c->psi_e[q]; }
    a.tag = c->d_tag; a.cls = c->d_cls; a.med = (const MedDev<T>*)c->d_med;
    a.dstate = (T*)c->disp_state;
    a.pstride = std::max(c->n_disp, 1);
    a.maxp = c->maxp;
    a.dcomp = (long)(2 + 2 * c->maxp) * a.pstride;
    a.disp0 = 0;
    a.dt_mu0 = (T)(c->d.dt / kMu0);
    a.dt_eps0 = (T)(c->d.dt / kEps0);
    a.nx = c->L.nx; a.ny = c->L.ny; a.nz = c->L.nz; a.sy = c->L.sy; a.sz = c->L.sz;
    a.sx = c->L.sx; a.off = c->L.off;
    a.npml = c->d.npml;
    a.cpml = c->cpml_mask;
    a.disp_tag0 = c->disp_tag0;
    a.xchunk = c->xchunk;
    a.ntz = (c->L.nz + kLanes - 1) / kLanes;
    a.nty = (c->L.ny + c->rows_y - 1) / c->rows_y;
    a.rows = c->rows_y;
    a.tile_order = c->tile_order;
    return a;
}

template <typename T>
unsigned yee_blocks(omx3d_ctx* c, const YeeArgs<T>& a) {
    const long ntx = (c->L.nx + c->xchunk - 1) / c->xchunk;
    return (unsigned)((long)a.ntz * a.nty * ntx);
}

template <typename T>
int wrap_e(omx3d_ctx* c, hipStream_t st) {
    const Layout& L = c->L;
    T** f = (T**)c->f;
    if (c->d.boundaries[0] == OMX_BC_PERIODIC)
        hipLaunchKernelGGL(k_wrap<T>, dim3(blocks_for(L.nz + 2, 256), L.ny + 2), dim3(256), 0, st, f[2], f[1], L, 0, L.nx + 1, 1);
    if (c->d.boundaries[1] == OMX_BC_PERIODIC)
        hipLaunchKernelGGL(k_wrap<T>, dim3(blocks_for(L.nz + 2, 256), L.nx + 2), dim3(256), 0, st, f[0], f[2], L, 1, L.ny + 1, 1);
    if (c->d.boundaries[2] == OMX_BC_PERIODIC)
        hipLaunchKernelGGL(k_wrap<T>, dim3(blocks_for(L.ny + 2, 256), L.nx + 2), dim3(256), 0, st, f[0], f[1], L, 2, L.nz + 1, 1);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

template <typename T>
int wrap_h(omx3d_ctx* c, hipStream_t st) {
    const Layout& L = c->L;
    T** f = (T**)c->f;
    if (c->d.boundaries[0] == OMX_BC_PERIODIC)
        hipLaunchKernelGGL(k_wrap<T>, dim3(blocks_for(L.nz + 2, 256), L.ny + 2), dim3(256), 0, st, f[4], f[5], L, 0, 0, L.nx);
    if (c->d.boundaries[1] == OMX_BC_PERIODIC)
        hipLaunchKernelGGL(k_wrap<T>, dim3(blocks_for(L.nz + 2, 256), L.nx + 2), dim3(256), 0, st, f[3], f[5], L, 1, 0, L.ny);
    if (c->d.boundaries[2] == OMX_BC_PERIODIC)
        hipLaunchKernelGGL(k_wrap<T>, dim3(blocks_for(L.ny + 2, 256), L.nx + 2), dim3(256), 0, st, f[3], f[4], L, 2, 0, L.nz);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

TfsfArgs tfsf_args(omx3d_ctx* c) {
    TfsfArgs t;
    t.i0 = c->tf_lo[0]; t.j0 = c->tf_lo[1]; t.k0 = c->tf_lo[2];
    t.i1 = c->tf_hi[0]; t.j1 = c->tf_hi[1]; t.k1 = c->tf_hi[2];
    t.kb = c->tf_kb; t.n1 = c->tf_n; t.ks = c->tf_ks;
    t.dt_mu0 = c->d.dt / kMu0; t.dt_eps0 = c->d.dt / kEps0; t.inv_dr = 1.0 / c->d.dr;
    return t;
}

double tfsf_source(const omx3d_ctx* c, double t_e) {
    const double x = (t_e - c->tf_t0) / c->tf_tau;
    return c->tf_e0 * std::exp(-(x * x)) * std::cos(c->tf_w0 * (t_e - c->tf_t0));
}

// Whether the TF/SF E corrections can ride in the E tiles: no cell of a
// dispersive class lies on the E correction faces (those cells are updated
// by the dispersive workgroups of the E launch, so their corrections stay in
// k_tfsf_e, after the launch).
int tfsf_fold_check(omx3d_ctx* c) {
    if (c->tf_fold_e >= 0) return OMX_OK;
    c->tf_fold_e = 0;
    if (!c->tfsf) return OMX_OK;
    if (c->media_set && c->n_classes > 1 && c->n_disp > 0) {
        int* d_flag = nullptr;
        int rc = dalloc(c, (void**)&d_flag, sizeof(int));
        if (rc) return rc;
        HIPCHK(hipMemsetAsync(d_flag, 0, sizeof(int), c->s));
        hipLaunchKernelGGL(k_disp_on_tfsf, dim3(blocks_for(c->n_disp, 256)), dim3(256), 0, c->s, c->d_disp,
                           c->n_disp, tfsf_args(c), d_flag);
        HIPCHK(hipGetLastError());
        int flag = 0;
        HIPCHK(hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        dfreet(d_flag);
        if (flag) return OMX_OK;
    }
    c->tf_fold_e = 1;
    return OMX_OK;
}

// TF/SF part of a separate H or E launch (workgroups of 64 x rows_y threads):
// the line's workgroups start at b0 (0 in the H launch).
TfsfLaunch tfsf_launch(omx3d_ctx* c, bool on) {
    TfsfLaunch L{};
    L.on = on ? 1 : 0;
    if (!on) return L;
    L.t = tfsf_args(c);
    L.ex1 = c->d_tf;
    L.hy1 = c->d_tf + c->tf_n;
    const int wg = kLanes * c->rows_y;
    L.nb = (c->tf_n + wg - 1) / wg;
    return L;
}

// Whether every array of the grid (fields, psi, tags, dispersive state) is
// addressed by 32-bit byte offsets: the field arrays are the largest.
bool offsets32(const omx3d_ctx* c) {
    return c->offset_bits == 32 && (unsigned long long)c->L.nalloc * c->esz < (1ull << 32);
}

// The H / E tile kernel instance for a grid: CPML, media, 32-bit offsets and
// TF/SF are template switches, so that a grid without one of them carries
// none of its code or registers.
template <typename T>
using H_kernel = void (*)(const YeeArgs<T>, const TfsfLaunch);
template <typename T>
using E_kernel = void (*)(const YeeArgs<T>, const long*, int, int, const TfsfLaunch);

template <typename T, bool CP, bool B32>
H_kernel<T> pick_h_tf(bool tf) {
    return tf ? k_h_update<T, CP, B32, true> : k_h_update<T, CP, B32, false>;
}
template <typename T>
H_kernel<T> pick_h(bool cp, bool b32, bool tf) {
    if (cp) return b32 ? pick_h_tf<T, true, true>(tf) : pick_h_tf<T, true, false>(tf);
    return b32 ? pick_h_tf<T, false, true>(tf) : pick_h_tf<T, false, false>(tf);
}
template <typename T, bool CP, int MED, bool B32>
E_kernel<T> pick_e_tf(bool tf) {
    return tf ? k_e_update<T, CP, MED, B32, true> : k_e_update<T, CP, MED, B32, false>;
}
template <typename T, bool CP, int MED>
E_kernel<T> pick_e_b32(bool b32, bool tf) {
    return b32 ? pick_e_tf<T, CP, MED, true>(tf) : pick_e_tf<T, CP, MED, false>(tf);
}
template <typename T>
E_kernel<T> pick_e(bool cp, bool md, bool b32, bool tf) {
    if (cp) return md ? pick_e_b32<T, true, 1>(b32, tf) : pick_e_b32<T, true, 0>(b32, tf);
    return md ? pick_e_b32<T, false, 1>(b32, tf) : pick_e_b32<T, false, 0>(b32, tf);
}

// ---------------------------------------------------------------------------
// Near-field patches (FP32 grids; kernels: k_patch).  The boxes follow the
// sources: each source's points widened by patch_radius cells on every side,
// clipped to the cells outside the CPML faces and off the outer planes
// (2..n-1), then merged while two boxes overlap or touch.
// ---------------------------------------------------------------------------
constexpr long kPatchAlign = 256;
inline long patch_padded(long n) { return (n + kPatchAlign - 1) / kPatchAlign * kPatchAlign; }

void free_patches(omx3d_ctx* c) {
    c->device_bytes -= c->patch_bytes;
    c->patch_bytes = 0;
    dfreet(c->d_boxes);
    dfreet(c->d_patch);
    dfreet(c->d_patch_cdt);
    dfreet(c->d_patch_dl);
    dfreet(c->d_patch_ds);
    for (auto& s : c->sources) {
        dfreet(s.d_sh);
        s.h_sh.assign(s.npts, nullptr);
    }
    for (int q = 0; q < 3; ++q) dfreet(c->d_aff_sh[q]);
    c->h_boxes.clear();
    c->patch_cells = 0;
}

// The shadow element of component comp of cell (i,j,k), or null.
double* patch_element(const omx3d_ctx* c, int comp, int i, int j, int k) {
    for (const PatchBox& B : c->h_boxes) {
        const int di = i - B.lo[0], dj = j - B.lo[1], dk = k - B.lo[2];
        if (di < 0 || dj < 0 || dk < 0 || di >= B.w[0] || dj >= B.w[1] || dk >= B.w[2]) continue;
        return B.pe + comp * B.cs + ((long)di * B.w[1] + dj) * B.w[2] + dk;
    }
    return nullptr;
}

template <typename P>
int upload_ptrs(omx3d_ctx* c, P** d, const std::vector<P>& h) {
    dfreet(*d);
    bool any = false;
    for (auto v : h) any = any || v != nullptr;
    if (!any) return OMX_OK;
    int rc = dalloc(c, (void**)d, sizeof(P) * h.size());
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(*d, h.data(), sizeof(P) * h.size(), hipMemcpyHostToDevice, c->s));
    return OMX_OK;
}

int build_patches_impl(omx3d_ctx* c) {
    if (c->prec != 4 || c->patch_radius <= 0 || c->tfsf || c->sources.empty()) return OMX_OK;
    const int n[3] = {c->d.nx, c->d.ny, c->d.nz};
    int amin[3], amax[3];
    for (int a = 0; a < 3; ++a) {
        const bool cp = c->d.boundaries[a] == OMX_BC_CPML;
        amin[a] = cp ? c->d.npml + 1 : 2;
        amax[a] = cp ? n[a] - c->d.npml : n[a] - 1;
    }
    struct Bx { int lo[3], hi[3]; };
    std::vector<Bx> bx;
    const int R = c->patch_radius;
    for (const auto& src : c->sources) {
        Bx b;
        for (int a = 0; a < 3; ++a) { b.lo[a] = INT32_MAX; b.hi[a] = INT32_MIN; }
        for (int p = 0; p < src.npts; ++p)
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = std::min(b.lo[a], src.h_ijk[3 * p + a]);
                b.hi[a] = std::max(b.hi[a], src.h_ijk[3 * p + a]);
            }
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = std::max(b.lo[a] - R, amin[a]);
            b.hi[a] = std::min(b.hi[a] + R, amax[a]);
            ok = ok && b.lo[a] <= b.hi[a];
        }
        if (ok) bx.push_back(b);
    }
    for (bool merged = true; merged;) {
        merged = false;
        for (size_t u = 0; u < bx.size() && !merged; ++u)
            for (size_t v = u + 1; v < bx.size() && !merged; ++v) {
                bool touch = true;
                for (int a = 0; a < 3; ++a)
                    touch = touch && bx[u].lo[a] <= bx[v].hi[a] + 1 && bx[v].lo[a] <= bx[u].hi[a] + 1;
                if (!touch) continue;
                for (int a = 0; a < 3; ++a) {
                    bx[u].lo[a] = std::min(bx[u].lo[a], bx[v].lo[a]);
                    bx[u].hi[a] = std::max(bx[u].hi[a], bx[v].hi[a]);
                }
                bx.erase(bx.begin() + v);
                merged = true;
            }
    }
    if (bx.empty()) return OMX_OK;
    // each box starts on a multiple of kPatchAlign cells (the patch kernels'
    // workgroup size), in the storage and in the launch's flat index
    long cells = 0;
    for (const Bx& b : bx) {
        const long n = (long)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) * (b.hi[2] - b.lo[2] + 1);
        if (n > INT32_MAX - kPatchAlign)   // the patch kernels index a box's cells in 32 bits
            return fail(OMX_ERR_RANGE, "near-field patch box of " + std::to_string(n) +
                                           " cells: lower the patch radius or set it to 0");
        cells += patch_padded(n);
    }
    int rc = dalloc(c, (void**)&c->d_patch, sizeof(double) * 6 * cells);
    if (rc) return rc;
    c->patch_cells = cells;
    // dt_eps0 / eps_r per class and component (dt_eps0 in vacuum), as the
    // FP64 engine's dielectric coefficient dt_eps0 * (1 / eps_r)
    const double dt_eps0 = c->d.dt / kEps0;
    const int ncls = c->media_set ? std::max(c->n_classes, 1) : 1;
    std::vector<int8_t> cls(3 * (size_t)ncls, 0);
    if (c->media_set && c->d_cls) HIPCHK(hipMemcpy(cls.data(), c->d_cls, cls.size(), hipMemcpyDeviceToHost));
    std::vector<double> cdt(3 * (size_t)ncls, dt_eps0);
    for (size_t q = 0; q < cls.size(); ++q) {
        const int m = cls[q];
        if (m > 0 && c->media[m].kind == OMX_MEDIUM_DIELECTRIC) cdt[q] = dt_eps0 * (1.0 / c->media[m].eps_r);
    }
    rc = dalloc(c, (void**)&c->d_patch_cdt, sizeof(double) * cdt.size());
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->d_patch_cdt, cdt.data(), sizeof(double) * cdt.size(), hipMemcpyHostToDevice, c->s));
    long first = 0;
    for (const Bx& b : bx) {
        PatchBox B;
        std::memset(&B, 0, sizeof(B));
        for (int a = 0; a < 3; ++a) { B.lo[a] = b.lo[a]; B.w[a] = b.hi[a] - b.lo[a] + 1; }
        B.first = first;
        B.pe = c->d_patch + first;
        B.ph = c->d_patch + 3 * cells + first;
        B.cs = cells;
        for (int q = 0; q < 6; ++q) B.den[q] = c->dcoef + c->co_den[q];
        B.cdt = c->d_patch_cdt;
        B.dt_mu0 = c->d.dt / kMu0;
        B.dt_eps0 = dt_eps0;
        B.g = 0;
        c->h_boxes.push_back(B);
        first += patch_padded((long)B.w[0] * B.w[1] * B.w[2]);
    }
    // cells of Drude / Drude-Lorentz classes in the boxes: their position in
    // the dispersive list, and an FP64 copy of their ADE state
    if (c->media_set && c->n_disp > 0 && (c->has_drude || c->has_dl)) {
        std::vector<long> list(c->n_disp);
        HIPCHK(hipMemcpy(list.data(), c->d_disp, sizeof(long) * list.size(), hipMemcpyDeviceToHost));
        std::vector<int> dl(cells, -1);
        bool any = false;
        for (int p = 0; p < c->n_disp; ++p) {
            const long e = list[p];
            const int i = (int)(e & 0xffff), j = (int)((e >> 16) & 0xffff), k = (int)((e >> 32) & 0xffff);
            for (const PatchBox& B : c->h_boxes) {
                const int di = i - B.lo[0], dj = j - B.lo[1], dk = k - B.lo[2];
                if (di < 0 || dj < 0 || dk < 0 || di >= B.w[0] || dj >= B.w[1] || dk >= B.w[2]) continue;
                dl[B.first + ((long)di * B.w[1] + dj) * B.w[2] + dk] = p;
                any = true;
            }
        }
        if (any) {
            const int nv = 2 + 2 * c->maxp;
            rc = dalloc(c, (void**)&c->d_patch_dl, sizeof(int) * cells);
            if (!rc) rc = dalloc(c, (void**)&c->d_patch_ds, sizeof(double) * 3 * nv * cells);
            if (rc) return rc;
            HIPCHK(hipMemcpyAsync(c->d_patch_dl, dl.data(), sizeof(int) * cells, hipMemcpyHostToDevice, c->s));
            for (PatchBox& B : c->h_boxes) {
                B.dl = c->d_patch_dl + B.first;
                B.ds = c->d_patch_ds + B.first;
                B.nv = nv;
                B.mco = c->d_med_co;
            }
            HIPCHK(hipStreamSynchronize(c->s));
        }
    }
    rc = dalloc(c, (void**)&c->d_boxes, sizeof(PatchBox) * c->h_boxes.size());
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->d_boxes, c->h_boxes.data(), sizeof(PatchBox) * c->h_boxes.size(),
                          hipMemcpyHostToDevice, c->s));
    for (auto& src : c->sources) {
        for (int p = 0; p < src.npts; ++p)
            src.h_sh[p] = patch_element(c, src.dir - 1, src.h_ijk[3 * p], src.h_ijk[3 * p + 1], src.h_ijk[3 * p + 2]);
        rc = upload_ptrs(c, &src.d_sh, src.h_sh);
        if (rc) return rc;
    }
    for (int comp = 0; comp < 3; ++comp) {
        const std::vector<int>& v = c->h_aff_ijk[comp];
        std::vector<double*> sh(v.size() / 3);
        for (size_t q = 0; q < sh.size(); ++q) sh[q] = patch_element(c, comp, v[3 * q], v[3 * q + 1], v[3 * q + 2]);
        rc = upload_ptrs(c, &c->d_aff_sh[comp], sh);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->s));
    c->patch_dirty = true;
    return OMX_OK;
}

int build_patches(omx3d_ctx* c) {
    free_patches(c);
    c->patches_built = true;
    c->patch_dirty = false;
    const long before = c->device_bytes;
    const int rc = build_patches_impl(c);
    c->patch_bytes = c->device_bytes - before;
    return rc;
}

template <typename T>
int launch_patch(omx3d_ctx* c, const YeeArgs<T>& a, int phase, hipStream_t st) {
    if (c->patch_cells == 0) return OMX_OK;
    hipLaunchKernelGGL(k_patch<T>, dim3(blocks_for(c->patch_cells, 256)), dim3(256), 0, st, a, c->d_boxes,
                       (int)c->h_boxes.size(), c->patch_cells, phase);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

// Boxes up to date with the sources, and shadows equal to the fields after
// a set-up or a write from the host.
template <typename T>
int ensure_patches(omx3d_ctx* c) {
    if (!c->patches_built) {
        int rc = build_patches(c);
        if (rc) return rc;
    }
    if (c->patch_dirty) {
        int rc = launch_patch<T>(c, make_args<T>(c), 2, c->s);
        if (rc) return rc;
        c->patch_dirty = false;
    }
    return OMX_OK;
}

template <typename T>
// tf: the TF/SF H corrections and the 1D line's H half step ride in the
// launch.  ev0/ev1 (profiling, may be null) bracket the H kernel itself.
int launch_h(omx3d_ctx* c, bool tf, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
    int rc = ensure_patches<T>(c);
    if (rc) return rc;
    rc = wrap_e<T>(c, c->s);
    if (rc) return rc;
    YeeArgs<T> a = make_args<T>(c);
    const dim3 blk(kLanes, c->rows_y);
    const TfsfLaunch L = tfsf_launch(c, tf);
    const unsigned nb = yee_blocks(c, a) + (L.on ? L.nb : 0);
    const bool cp = c->cpml_mask != 0, b32 = offsets32(c);
    if (ev0) HIPCHK(hipEventRecord(ev0, c->s));
    const H_kernel<T> k = pick_h<T>(cp, b32, L.on != 0);
    hipLaunchKernelGGL(k, dim3(nb), blk, 0, c->s, a, L);
    if (ev1) HIPCHK(hipEventRecord(ev1, c->s));
    HIPCHK(hipGetLastError());
    return launch_patch<T>(c, a, 0, c->s);
}

template <typename T>
int launch_j_correct(omx3d_ctx* c, double tau) {
    if (c->nq == 0) return OMX_OK;
    const double dt_eps0 = c->d.dt / kEps0;
    for (int comp = 0; comp < 3; ++comp) {
        if (c->n_aff[comp] == 0) continue;
        hipLaunchKernelGGL(k_j_correct<T>, dim3(blocks_for(c->n_aff[comp], 256)), dim3(256), 0, c->s,
                           (T*)c->f[comp], c->d_aff_off[comp], c->d_aff_a[comp], c->d_aff_b[comp],
                           c->n_aff[comp], comp, c->d_J, tau, dt_eps0,
                           c->media_set ? c->d_tag : (const uint8_t*)nullptr, c->d_cls, c->d_kinds,
                           (double* const*)c->d_aff_sh[comp]);
    }
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

// E of the cells of dispersive classes: a's E arrays are the input, the
// output arrays get the result (the same arrays for the separate kernels).
template <typename T>
int launch_e_dispersive(omx3d_ctx* c, const YeeArgs<T>& a, T* exo, T* eyo, T* ezo, hipStream_t st) {
    const unsigned nb = blocks_for(c->n_disp, 256);
    const bool b32 = offsets32(c);
    if (c->cpml_mask && b32)
        hipLaunchKernelGGL((k_e_dispersive<T, true, true>), dim3(nb), dim3(256), 0, st, a, exo, eyo, ezo, c->d_disp,
                           c->n_disp);
    else if (c->cpml_mask)
        hipLaunchKernelGGL((k_e_dispersive<T, true, false>), dim3(nb), dim3(256), 0, st, a, exo, eyo, ezo,
                           c->d_disp, c->n_disp);
    else if (b32)
        hipLaunchKernelGGL((k_e_dispersive<T, false, true>), dim3(nb), dim3(256), 0, st, a, exo, eyo, ezo,
                           c->d_disp, c->n_disp);
    else
        hipLaunchKernelGGL((k_e_dispersive<T, false, false>), dim3(nb), dim3(256), 0, st, a, exo, eyo, ezo,
                           c->d_disp, c->n_disp);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

// The E tiles and, in the same launch, the cells of dispersive classes; tf:
// the TF/SF E corrections and the 1D line's E half step (source at time
// t_tf) ride in the launch too.  ev0/ev1 (profiling, may be null) bracket
// that kernel.
template <typename T>
int launch_e(omx3d_ctx* c, double time, double t_skip, bool tf, double t_tf, hipEvent_t ev0 = nullptr,
             hipEvent_t ev1 = nullptr) {
    int rc = ensure_patches<T>(c);
    if (rc) return rc;
    rc = wrap_h<T>(c, c->s);
    if (rc) return rc;
    YeeArgs<T> a = make_args<T>(c);
    const dim3 blk(kLanes, c->rows_y);
    const bool cp = c->cpml_mask != 0, md = c->media_set && c->n_classes > 1;
    const int nd = md ? c->n_disp : 0;
    const int db = md ? (nd + kLanes * c->rows_y - 1) / (kLanes * c->rows_y) : 0;
    TfsfLaunch L = tfsf_launch(c, tf);
    L.b0 = db;
    if (tf) L.src = tfsf_source(c, t_tf);
    const dim3 grid(yee_blocks(c, a) + db + (L.on ? L.nb : 0));
    if (ev0) HIPCHK(hipEventRecord(ev0, c->s));
    const E_kernel<T> k = pick_e<T>(cp, md, offsets32(c), L.on != 0);
    hipLaunchKernelGGL(k, grid, blk, 0, c->s, a, (const long*)c->d_disp, nd, db, L);
    if (ev1) HIPCHK(hipEventRecord(ev1, c->s));
    HIPCHK(hipGetLastError());
    rc = launch_patch<T>(c, a, 1, c->s);
    if (rc) return rc;
    return launch_j_correct<T>(c, time - t_skip);
}

SourceParams source_params(const omx3d_ctx* c, const SourceDev& s) {
    SourceParams p;
    p.j_amp = s.j_amp; p.w0 = s.w0; p.t0 = s.t0; p.tau = s.tau; p.t_init = s.t_init; p.t_final = s.t_final;
    p.phase = s.phase;
    p.c_src = (((c->d.dt / kEps0) / c->d.dr) / kC0) / 2.0;  // interactions_mod.f90:32
    p.dt = c->d.dt;
    return p;
}

template <typename T>
int launch_sources(omx3d_ctx* c, double time) {
    int rc = ensure_patches<T>(c);
    if (rc) return rc;
    for (auto& s : c->sources) {
        if (!(time >= s.t_init && time <= s.t_final)) continue;  // J_mat = 0 outside the window
        hipLaunchKernelGGL(k_source<T>, dim3(blocks_for(s.npts, 256)), dim3(256), 0, c->s,
                           (T*)c->f[s.dir - 1], s.d_off, s.d_ka, s.d_kb, s.npts, source_params(c, s), time,
                           (double* const*)s.d_sh);
    }
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

template <typename T>
int launch_tfsf_e(omx3d_ctx* c, double t_e) {
    if (!c->tfsf) return OMX_OK;
    TfsfArgs t = tfsf_args(c);
    const long nxi = t.i1 - t.i0, nyj = t.j1 - t.j0 + 1, nzk = t.k1 - t.k0;
    const long n = 2 * nxi * nyj + 2 * nyj * nzk + t.n1;
    const double src = tfsf_source(c, t_e);
    hipLaunchKernelGGL(k_tfsf_e<T>, dim3(blocks_for(n, 256)), dim3(256), 0, c->s, make_args<T>(c), t,
                       c->d_tf, c->d_tf + c->tf_n, src);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

template <typename T>
int launch_record(omx3d_ctx* c, double* out) {
    if (c->n_samples == 0) return OMX_OK;
    hipLaunchKernelGGL(k_sample<T>, dim3(blocks_for(c->n_samples, 256)), dim3(256), 0, c->s, make_args<T>(c),
                       c->d_det_field, c->d_det_off, c->n_samples, out);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

int upload_detectors(omx3d_ctx* c) {
    if (!c->det_dirty) return OMX_OK;
    dfreet(c->d_det_field);
    dfreet(c->d_det_off);
    if (c->n_samples > 0) {
        int rc = dalloc(c, (void**)&c->d_det_field, sizeof(int) * c->n_samples);
        if (rc) return rc;
        rc = dalloc(c, (void**)&c->d_det_off, sizeof(long) * c->n_samples);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(c->d_det_field, c->h_det_field.data(), sizeof(int) * c->n_samples, hipMemcpyHostToDevice, c->s));
        HIPCHK(hipMemcpyAsync(c->d_det_off, c->h_det_off.data(), sizeof(long) * c->n_samples, hipMemcpyHostToDevice, c->s));
    }
    c->det_dirty = false;
    return OMX_OK;
}

template <typename T>
int launch_two_level(omx3d_ctx* c) {
    if (c->nq == 0 || !c->tls) return OMX_OK;
    hipLaunchKernelGGL(k_gather_e<T>, dim3(blocks_for(c->nq, 256)), dim3(256), 0, c->s, make_args<T>(c),
                       c->d_q_off, c->nq, c->d_q_e);
    hipLaunchKernelGGL(k_tls_step, dim3(blocks_for(c->nq, 256)), dim3(256), 0, c->s, c->d_tls, c->d_q_e, c->nq,
                       c->q_dt, c->d_q_dpdt);
    hipLaunchKernelGGL(k_scatter_j, dim3(blocks_for(c->nslot, 256)), dim3(256), 0, c->s, c->d_q_slot, c->nslot,
                       c->d_q_dpdt, c->q_dt, c->d_J);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

int ensure_events(omx3d_ctx* c, size_t n) {
    while (c->ev.size() < n) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->ev.push_back(e);
    }
    return OMX_OK;
}

size_t field_bytes(const omx3d_ctx* c) { return c->esz * (size_t)c->L.nalloc; }

// Sizes of the H psi arrays per axis (x, y, z slabs).
size_t psi_h_elems(const omx3d_ctx* c, int axis) {
    const size_t A = 128 / c->esz;
    if (axis == 0) return c->nsh * c->L.sx + A;
    if (axis == 1) return (size_t)(c->d.nx + 2) * c->nsh * c->L.sz + A;
    return (size_t)zpsi_elems(c->L.sy, c->d.nx, c->d.npml);
}

// Second field / H-psi buffers of the fused step: a copy of the current ones
// (boundary values never change, interior values are rewritten each step).
int ensure_alt(omx3d_ctx* c) {
    if (c->f_alt[0]) return OMX_OK;
    for (int q = 0; q < 6; ++q) {
        int rc = dalloc(c, &c->f_alt[q], field_bytes(c));
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(c->f_alt[q], c->f[q], field_bytes(c), hipMemcpyDeviceToDevice, c->s));
    }
    for (int q = 0; q < 6; ++q) {
        if (!c->psi_h[q]) continue;
        const size_t b = c->esz * psi_h_elems(c, q / 2);
        int rc = dalloc(c, &c->psi_h_alt[q], b);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(c->psi_h_alt[q], c->psi_h[q], b, hipMemcpyDeviceToDevice, c->s));
    }
    return OMX_OK;
}

// Back to the state before the first step: fields (both buffer sets), CPML
// psi (both H sets), dispersive state and the TF/SF line at zero, near-field
// patches rebuilt from the fields at the next launch (the tuners time steps
// on the grid itself and leave it as they found it, before a run).
int reset_run_state(omx3d_ctx* c) {
    const size_t fb = c->esz * c->L.nalloc;
    for (auto p : c->f) HIPCHK(hipMemsetAsync(p, 0, fb, c->s));
    for (auto p : c->f_alt) if (p) HIPCHK(hipMemsetAsync(p, 0, fb, c->s));
    const size_t A = 128 / c->esz;
    const size_t xs[2] = {c->nsh * c->L.sx + A, c->nse * c->L.sx + A};
    const size_t ys[2] = {(size_t)(c->d.nx + 2) * c->nsh * c->L.sz + A, (size_t)(c->d.nx + 2) * c->nse * c->L.sz + A};
    const size_t zs[2] = {(size_t)zpsi_elems(c->L.sy, c->d.nx, c->d.npml), (size_t)zpsi_elems(c->L.sy, c->d.nx, c->d.npml)};
    for (int a = 0; a < 3; ++a) {
        const size_t* sz = a == 0 ? xs : (a == 1 ? ys : zs);
        for (int q = 0; q < 2; ++q) {
            if (c->psi_h[2 * a + q]) HIPCHK(hipMemsetAsync(c->psi_h[2 * a + q], 0, c->esz * sz[0], c->s));
            if (c->psi_h_alt[2 * a + q]) HIPCHK(hipMemsetAsync(c->psi_h_alt[2 * a + q], 0, c->esz * sz[0], c->s));
            if (c->psi_e[2 * a + q]) HIPCHK(hipMemsetAsync(c->psi_e[2 * a + q], 0, c->esz * sz[1], c->s));
        }
    }
    if (c->disp_state) HIPCHK(hipMemsetAsync(c->disp_state, 0, disp_state_bytes(c), c->s));
    if (c->d_tf) HIPCHK(hipMemsetAsync(c->d_tf, 0, sizeof(double) * 2 * (size_t)c->tf_n, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->patch_dirty = true;
    return OMX_OK;
}

// Whether a source point lies in a Drude/DL component (the fused step adds
// sources after the E update, which equals the reference order only where
// the E update is linear in E without memory, i.e. vacuum and dielectric).
int check_sources_dispersive(omx3d_ctx* c) {
    if (c->src_in_dispersive >= 0) return OMX_OK;
    c->src_in_dispersive = 0;
    if (!c->media_set || c->n_classes <= 1 || c->sources.empty()) return OMX_OK;
    std::vector<int8_t> cls(3 * (size_t)c->n_classes);
    HIPCHK(hipMemcpy(cls.data(), c->d_cls, cls.size(), hipMemcpyDeviceToHost));
    for (auto& s : c->sources) {
        std::vector<long> off(s.npts);
        HIPCHK(hipMemcpy(off.data(), s.d_off, sizeof(long) * s.npts, hipMemcpyDeviceToHost));
        for (long o : off) {
            uint8_t tg = 0;
            HIPCHK(hipMemcpy(&tg, c->d_tag + o, 1, hipMemcpyDeviceToHost));
            const int m = cls[3 * tg + s.dir - 1];
            if (m && (c->media[m].kind == OMX_MEDIUM_DRUDE || c->media[m].kind == OMX_MEDIUM_DL)) {
                c->src_in_dispersive = 1;
                return OMX_OK;
            }
        }
    }
    return OMX_OK;
}

// Wavefronts along z and y of a k_step_wg workgroup for this grid (nzw = 0:
// the z extent does not fit one workgroup).  FP32 takes up to 16 wavefronts,
// FP64 up to 8 (LDS of the hand-overs, 64 KiB at most).
void wg_shape(const omx3d_ctx* c, int& nzw, int& nyw) {
    nzw = (c->d.nz + 63) / 64;
    const int maxw = c->prec == 4 ? 16 : 8;
    if (nzw > maxw) { nzw = 0; nyw = 0; return; }
    nyw = c->wg_ny > 0 ? c->wg_ny : std::max(1, maxw / nzw);
    nyw = std::max(1, std::min(nyw, maxw / nzw));
}

bool fused_applies(omx3d_ctx* c) {
    if (c->step_mode != 0) return false;
    int nzw, nyw;
    wg_shape(c, nzw, nyw);
    if (nzw == 0) return false;
    for (int a = 0; a < 3; ++a)
        if (c->d.boundaries[a] == OMX_BC_PERIODIC) return false;
    // the fused kernel addresses field and z-face psi arrays by 32-bit byte
    // offsets, and handles one z face per cell (the faces do not meet)
    if ((unsigned long long)c->L.nalloc * c->esz >= (1ull << 32)) return false;
    for (int a = 0; a < 3; ++a) {
        // E psi arrays have npml slots per side to the npml-1 of H
        const unsigned long long e = a == 2 ? 0ull : (unsigned long long)(c->nse - c->nsh) * (a == 0 ? c->L.sx : (c->d.nx + 2) * c->L.sz);
        if (((unsigned long long)psi_h_elems(c, a) + e) * c->esz >= (1ull << 32)) return false;
    }
    if ((c->cpml_mask & 48u) && c->d.nz < 2 * c->d.npml + 1) return false;
    return c->src_in_dispersive == 0;
}

template <typename T, int R, int CP>
void launch_wg_kernel(omx3d_ctx* c, bool md, bool tf, const YeeArgs<T>& a, const FusedOut<T>& fo, const WgPlan& P,
                      hipStream_t st) {
    const size_t lds = wg_lds_bytes(P.nzw * P.nyw, R, sizeof(T), tf);
    const dim3 blk(64 * P.nzw * P.nyw);
    // with TF/SF one workgroup more: the incident line's H half step
    const unsigned nb = (unsigned)P.box[P.nbox - 1].b_end + (tf ? 1u : 0u);
#define OMX_WG(MD, TF) hipLaunchKernelGGL((k_step_wg<T, R, CP, MD, TF>), dim3(nb), blk, lds, st, a, fo, P)
    if (md && tf) OMX_WG(1, true);
    else if (md) OMX_WG(1, false);
    else if (tf) OMX_WG(0, true);
    else OMX_WG(0, false);
#undef OMX_WG
}

template <typename T, int CP>
void launch_wg_rows(omx3d_ctx* c, bool md, bool tf, const YeeArgs<T>& a, const FusedOut<T>& fo, const WgPlan& P,
                    hipStream_t st) {
    if (c->wg_rows == 2) launch_wg_kernel<T, 2, CP>(c, md, tf, a, fo, P, st);
    else launch_wg_kernel<T, 1, CP>(c, md, tf, a, fo, P, st);
}

// Boxes of the fused step's launch, heaviest first: the x slabs holding the
// x CPML faces (planes 1..npml, nx-npml+1..nx; chunks of 4 planes), the y
// slabs (whole row blocks covering rows 1..npml and ny-npml+1..ny of the
// other planes; half chunks), then the interior (full chunks).  A slab tile
// carries the x/y psi, up to twice the bytes of an interior one; listed
// first, the slab tiles start first and the interior ones fill in behind
// them.  Without x or y CPML faces the interior is the whole grid.
WgPlan wg_plan(const omx3d_ctx* c) {
    WgPlan P;
    std::memset(&P, 0, sizeof(P));
    wg_shape(c, P.nzw, P.nyw);
    const int R = c->wg_rows, H = P.nyw * R;
    const int nx = c->d.nx, ny = c->d.ny, np = c->d.npml;
    const unsigned m = c->cpml_mask;
    const int X = std::max(1, c->wg_xchunk);
    auto add = [&](int i_lo, int i_hi, int j_lo, int j_hi, int xc) {
        if (i_lo > i_hi || j_lo > j_hi) return;
        WgBox& B = P.box[P.nbox];
        const int prev = P.nbox ? P.box[P.nbox - 1].b_end : 0;
        B.i_lo = i_lo; B.i_hi = i_hi; B.j_lo = j_lo; B.j_hi = j_hi;
        B.xc = std::max(1, std::min(xc, i_hi - i_lo + 1));
        B.nyb = (j_hi - j_lo + H) / H;
        B.b_end = prev + B.nyb * ((i_hi - i_lo + B.xc) / B.xc);
        ++P.nbox;
    };
    const int xa = (m & 1u) ? np + 1 : 1, xb = (m & 2u) ? nx - np : nx;
    const int hs = (np + H - 1) / H * H;
    const int ya = (m & 4u) ? hs + 1 : 1, yb = (m & 8u) ? ny - hs : ny;
    if (xa > xb || ya > yb) {
        add(1, nx, 1, ny, 4);
        return P;
    }
    add(1, xa - 1, 1, ny, 4);
    add(xb + 1, nx, 1, ny, 4);
    add(xa, xb, 1, ya - 1, std::max(1, X / 2));
    add(xa, xb, yb + 1, ny, std::max(1, X / 2));
    add(xa, xb, ya, yb, X);
    return P;
}

// Dispersive cells up to which the post kernel of the fused step loads each
// cell's whole state first (about one round of threads over the chip).
constexpr int kPostLoadsFirstMax = 1 << 19;

// Whether the fused kernel must tell media apart (its MED variant: class
// tags, the dielectric update, no store for cells of dispersive classes).
// Not when every class besides vacuum has a Drude / Drude-Lorentz component:
// those cells are all recomputed by the dispersive pass after the kernel,
// which overwrites the vacuum update the kernel stored; on CPML faces the
// kernel has then already advanced their E psi (a recursion in H only), and
// the pass adds those values instead of stepping them again.
int wg_media_check(omx3d_ctx* c) {
    if (c->wg_med >= 0) return OMX_OK;
    c->wg_med = (!c->media_set || c->n_classes <= 1 || (c->disp_tag0 == 1 && c->n_disp > 0)) ? 0 : 1;
    return OMX_OK;
}

// The fused step: one launch over every box, with the CPML code of the
// faces the grid has compiled in (the interior tiles skip the x/y psi code
// by uniform branches).  Separate launches for slabs and interior, in turn
// or on two streams, measured slower: the slab launch alone cannot fill the
// chip, and the interior tiles then run without slab tiles to balance them.
template <typename T>
int launch_wg(omx3d_ctx* c, bool md, bool tf, const YeeArgs<T>& a, const FusedOut<T>& fo) {
    const WgPlan P = wg_plan(c);
    if (P.nzw == 0) return fail(OMX_ERR_STATE, "fused step: the z extent does not fit one workgroup");
    int rc = wg_media_check(c);
    if (rc) return rc;
    md = md && c->wg_med == 1;
    if (c->cpml_mask & 15u) launch_wg_rows<T, 7>(c, md, tf, a, fo, P, c->s);
    else if (c->cpml_mask & 48u) launch_wg_rows<T, 4>(c, md, tf, a, fo, P, c->s);
    else launch_wg_rows<T, 0>(c, md, tf, a, fo, P, c->s);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

// One fused step (k_step_wg): reads the current buffers, writes the other
// set, then swaps them; the cells of dispersive classes (with the TF/SF
// line's E half step) and the near-field patches follow in their own
// launches.
template <typename T>
int launch_fused(omx3d_ctx* c, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, bool fold_e = false,
                 double t_tf = 0.0) {
    int rc = ensure_patches<T>(c);
    if (rc) return rc;
    rc = ensure_alt(c);
    if (rc) return rc;
    const YeeArgs<T> a = make_args<T>(c);
    FusedOut<T> fo;
    std::memset(&fo, 0, sizeof(fo));
    fo.ex = (T*)c->f_alt[0]; fo.ey = (T*)c->f_alt[1]; fo.ez = (T*)c->f_alt[2];
    fo.hx = (T*)c->f_alt[3]; fo.hy = (T*)c->f_alt[4]; fo.hz = (T*)c->f_alt[5];
    for (int q = 0; q < 6; ++q) fo.ph[q] = (T*)c->psi_h_alt[q];
    // TF/SF: the kernel reads the incident line's E and advances its H in
    // an extra workgroup
    if (c->tfsf) { fo.ex1 = c->d_tf; fo.hy1 = c->d_tf + c->tf_n; fo.t = tfsf_args(c); }
    const bool md = c->media_set && c->n_classes > 1, tf = c->tfsf;
    if (ev0) HIPCHK(hipEventRecord(ev0, c->s));
    rc = launch_wg<T>(c, md, tf, a, fo);
    if (rc) return rc;
    if (ev1) HIPCHK(hipEventRecord(ev1, c->s));
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 6; ++q) std::swap(c->f[q], c->f_alt[q]);
    for (int q = 0; q < 6; ++q) std::swap(c->psi_h[q], c->psi_h_alt[q]);
    if ((md && c->n_disp > 0) || (tf && fold_e)) {
        // one launch for the cells of dispersive classes (H^{n+1/2} now
        // current, E^n in the other buffer, E^{n+1} into the current one) and
        // the TF/SF E step (corrections to E^{n+1}, the line's E half step)
        YeeArgs<T> an = make_args<T>(c);
        an.ex = (T*)c->f_alt[0]; an.ey = (T*)c->f_alt[1]; an.ez = (T*)c->f_alt[2];
        const int nd = md ? (int)blocks_for(c->n_disp, 256) : 0;
        const TfsfArgs t = tf ? tfsf_args(c) : TfsfArgs{};
        const long n_tf = 2L * (t.i1 - t.i0) * (t.j1 - t.j0 + 1) + 2L * (t.j1 - t.j0 + 1) * (t.k1 - t.k0) + t.n1;
        const int nl = tf && fold_e ? (int)blocks_for(n_tf, 256) : 0;
        const double src = tf && fold_e ? tfsf_source(c, t_tf) : 0.0;
        double* ex1 = tf ? c->d_tf : nullptr;
        const double* hy1 = tf ? c->d_tf + c->tf_n : nullptr;
        const int n = md ? c->n_disp : 0;
        T *exo = (T*)c->f[0], *eyo = (T*)c->f[1], *ezo = (T*)c->f[2];
        // a list that one round of threads covers is latency-bound: its cells
        // load everything first (measured: the 256^3 grid's 24280 sphere cells
        // 13.3 -> 10.9 us; a 4.4 M-cell slab 374 -> 380 us, so not there)
        const bool lf = n <= kPostLoadsFirstMax;
#define OMX_POST(CP, ST, LF)                                                                                     \
    hipLaunchKernelGGL((k_wg_post<T, CP, true, ST, LF>), dim3(nd + nl), dim3(256), 0, c->s, an, exo, eyo, ezo, \
                       c->d_disp, n, nd, t, ex1, hy1, src)
        if (c->cpml_mask && c->wg_med == 0) {
            if (lf) OMX_POST(true, true, true);
            else OMX_POST(true, true, false);
        } else if (c->cpml_mask) {
            if (lf) OMX_POST(true, false, true);
            else OMX_POST(true, false, false);
        } else {
            if (lf) OMX_POST(false, false, true);
            else OMX_POST(false, false, false);
        }
#undef OMX_POST
        HIPCHK(hipGetLastError());
    }
    if (c->patch_cells) {
        // near-field patches: H^{n+1/2} from E^n (the other buffer) into the
        // current H, then E^{n+1} from it into the current E
        YeeArgs<T> ah = make_args<T>(c);
        ah.ex = (T*)c->f_alt[0]; ah.ey = (T*)c->f_alt[1]; ah.ez = (T*)c->f_alt[2];
        rc = launch_patch<T>(c, ah, 0, c->s);
        if (rc) return rc;
        rc = launch_patch<T>(c, make_args<T>(c), 1, c->s);
        if (rc) return rc;
    }
    return OMX_OK;
}

template <typename T>
int run_impl(omx3d_ctx* c, int tt_first, int n_steps, int record_every) {
    int rc = upload_detectors(c);
    if (rc) return rc;
    const double dt = c->d.dt;
    size_t evi = 0;
    const int pe = c->prof_every;
    if (pe > 0) {
        rc = ensure_events(c, (size_t)4 * ((n_steps + pe - 1) / pe));
        if (rc) return rc;
    }
    rc = check_sources_dispersive(c);
    if (rc) return rc;
    const bool fused = fused_applies(c);
    c->last_fused = fused;
    // TF/SF inside the separate H and E launches (the E side only when no
    // cell of a dispersive class lies on the E correction faces: the tiles
    // leave such cells to the dispersive workgroups)
    rc = tfsf_fold_check(c);
    if (rc) return rc;
    const bool fold = !fused && c->tfsf, fold_e = fold && c->tf_fold_e == 1;
    // the fused step's post kernel adds the TF/SF E corrections and advances
    // the line's E (same condition as fold_e)
    const bool fold_wg = fused && c->tfsf && c->tf_fold_e == 1;
    const int tt_0 = c->n_skip == 1 ? 1 : 0;
    for (int tt = tt_first; tt < tt_first + n_steps; ++tt) {
        const double time = dt * ((double)tt - 0.5);
        // a new quantum step starts when tq_step = 1 + int((tt - tt_0)/n_skip)
        // grows: t_skip = time and the quantum systems move this step
        // (mxim_mxll.f90:76-77,97-104); with n_skip = 1 that is every step
        const int tq_step = 1 + (tt - tt_0) / c->n_skip;
        bool move = false;
        if (tq_step > c->tq_step_old) {
            c->t_skip = time;
            c->tq_step_old = tq_step;
            move = c->nq > 0;
        }
        const double t_skip = c->t_skip;
        // profiling: events bracket the H and E kernels themselves (the
        // fused kernel takes the E slot, the H slot stays empty)
        const bool prof = pe > 0 && (tt - tt_first) % pe == 0;
        hipEvent_t* ev = prof ? &c->ev[evi] : nullptr;
        if (prof) evi += 4;
        if (fused) {
            // H and E in one pass; sources then enter E after its update
            // (same value as before it in vacuum/dielectric source cells).
            if (ev) {
                HIPCHK(hipEventRecord(ev[0], c->s));
                HIPCHK(hipEventRecord(ev[1], c->s));
            }
            rc = launch_fused<T>(c, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr, fold_wg, dt * (double)tt);
            if (rc) return rc;
            rc = launch_sources<T>(c, time);
            if (rc) return rc;
            rc = launch_j_correct<T>(c, time - t_skip);
            if (rc) return rc;
        } else {
            rc = launch_h<T>(c, fold, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr);
            if (rc) return rc;
            rc = launch_sources<T>(c, time);
            if (rc) return rc;
            rc = launch_e<T>(c, time, t_skip, fold_e, dt * (double)tt, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr);
            if (rc) return rc;
        }
        if (!fold_e && !fold_wg) {
            rc = launch_tfsf_e<T>(c, dt * (double)tt);
            if (rc) return rc;
        }
        if (move) {
            rc = launch_two_level<T>(c);
            if (rc) return rc;
        }
        if (record_every > 0 && tt % record_every == 0 && c->n_samples > 0) {
            if (c->rec_rows >= c->rec_cap) return fail(OMX_ERR_RANGE, "detector record buffer full");
            rc = launch_record<T>(c, c->d_rec + (long)c->rec_rows * c->n_samples);
            if (rc) return rc;
            c->rec_rows++;
        }
    }
    if (pe > 0) {
        HIPCHK(hipStreamSynchronize(c->s));
        for (size_t q = 0; q + 3 < evi; q += 4) {
            float mh = 0.f, me = 0.f;
            HIPCHK(hipEventElapsedTime(&mh, c->ev[q], c->ev[q + 1]));
            HIPCHK(hipEventElapsedTime(&me, c->ev[q + 2], c->ev[q + 3]));
            c->prof_ms_h += mh;
            c->prof_ms_e += me;
            c->prof_n += 1;
        }
    }
    return OMX_OK;
}

// The energy reduction on the context's stream: partial sums of
// kEnergyBlocks workgroups into part (3 * kEnergyBlocks doubles), the result
// (w64, w32c) into out (2 doubles, device memory).
template <typename T>
int launch_energy(omx3d_ctx* c, double* part, double* out) {
    const int nb = kEnergyBlocks;
    YeeArgs<T> a = make_args<T>(c);
    const bool md = c->media_set && c->n_classes > 1;
    if (md) hipLaunchKernelGGL((k_energy<T, true>), dim3(nb), dim3(256), 0, c->s, a, c->d_eps_class, part);
    else hipLaunchKernelGGL((k_energy<T, false>), dim3(nb), dim3(256), 0, c->s, a, c->d_eps_class, part);
    const double dr = c->d.dr;
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, c->s, part, nb, dr * dr * dr, out);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

template <typename T>
int energy_impl(omx3d_ctx* c, double* w64, double* w32c) {
    const int nb = kEnergyBlocks;
    if (!c->d_red) {
        int rc = dalloc(c, (void**)&c->d_red, sizeof(double) * (3 * nb + 2));
        if (rc) return rc;
    }
    int rc = launch_energy<T>(c, c->d_red, c->d_red + 3 * nb);
    if (rc) return rc;
    double out[2];
    HIPCHK(hipMemcpyAsync(out, c->d_red + 3 * nb, sizeof(out), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (w64) *w64 = out[0];
    if (w32c) *w32c = out[1];
    return OMX_OK;
}

// Device address of `scratch` on each GPU, taken once per device with that
// device current (each GPU holds its own copy of the array).
constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_on(int dev, double** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(OMX_ERR_RANGE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OMX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return OMX_OK;
}

// Field bounds of the reference arrays (mxll_3D_mod.f90:123-131).
void field_bounds(const Layout& L, int field, int lo[3], int hi[3]) {
    const int n[3] = {L.nx, L.ny, L.nz};
    for (int a = 0; a < 3; ++a) { lo[a] = 1; hi[a] = n[a] + 1; }
    if (field <= 3) {
        const int ax = field - 1;       // E component: 0:n along its own axis
        lo[ax] = 0; hi[ax] = n[ax];
    } else {
        const int ax = field - 4;       // H component: 1:n along its axis, 0:n across
        for (int a = 0; a < 3; ++a) { lo[a] = 0; hi[a] = n[a]; }
        lo[ax] = 1; hi[ax] = n[ax];
    }
}

template <typename T>
int copy_field_impl(omx3d_ctx* c, int field, double* host, int to_device) {
    const Layout& L = c->L;
    std::vector<T> stage(L.nalloc);
    T* dptr = (T*)c->f[field - 1];
    HIPCHK(hipMemcpyAsync(stage.data(), dptr, sizeof(T) * L.nalloc, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    int lo[3], hi[3];
    field_bounds(L, field, lo, hi);
    const long e0 = hi[0] - lo[0] + 1, e1 = hi[1] - lo[1] + 1;
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
                const long h = (i - lo[0]) + e0 * ((j - lo[1]) + e1 * (k - lo[2]));
                if (to_device) stage[L.at(i, j, k)] = (T)host[h];
                else host[h] = (double)stage[L.at(i, j, k)];
            }
    if (to_device) {
        HIPCHK(hipMemcpyAsync(dptr, stage.data(), sizeof(T) * L.nalloc, hipMemcpyHostToDevice, c->s));
        if (c->f_alt[field - 1])   // the fused step's second buffer holds the same boundary values
            HIPCHK(hipMemcpyAsync(c->f_alt[field - 1], stage.data(), sizeof(T) * L.nalloc, hipMemcpyHostToDevice, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        c->patch_dirty = true;
    }
    return OMX_OK;
}

// Classes of per-component medium triples present in the grid; tag = class.
int build_tags(omx3d_ctx* c) {
    const Layout& L = c->L;
    const int nm = c->n_media;
    const int ncode = (nm + 1) * (nm + 1) * (nm + 1);
    const long ncell = (long)L.nx * L.ny * L.nz;
    unsigned* d_present = nullptr;
    uint8_t* d_lut = nullptr;
    int rc = dalloc(c, (void**)&d_present, sizeof(unsigned) * ((ncode + 31) / 32));
    if (rc) return rc;
    hipLaunchKernelGGL(k_media_codes, dim3(blocks_for(ncell, 256)), dim3(256), 0, c->s, c->d_aux, L.nx, L.ny, L.nz, nm,
                       d_present, (const uint8_t*)nullptr, (uint8_t*)nullptr, L, (int32_t*)nullptr);
    HIPCHK(hipGetLastError());
    std::vector<unsigned> present((ncode + 31) / 32);
    HIPCHK(hipMemcpyAsync(present.data(), d_present, sizeof(unsigned) * present.size(), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    std::vector<uint8_t> lut(ncode, 0);
    std::vector<int8_t> cls = {0, 0, 0};
    std::vector<double> epsc = {1.0, 1.0, 1.0};
    int ncls = 1;
    // classes without a Drude / Drude-Lorentz component first, so that the
    // dispersive ones are the tags >= disp_tag0
    c->disp_tag0 = 256;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) c->disp_tag0 = ncls;
        for (int code = 1; code < ncode; ++code) {
            if (!((present[code >> 5] >> (code & 31)) & 1u)) continue;
            const int m0 = code % (nm + 1), m1 = (code / (nm + 1)) % (nm + 1), m2 = code / ((nm + 1) * (nm + 1));
            const int mm[3] = {m0, m1, m2};
            bool disp = false;
            for (int q = 0; q < 3; ++q)
                if (mm[q] && (c->media[mm[q]].kind == OMX_MEDIUM_DRUDE || c->media[mm[q]].kind == OMX_MEDIUM_DL))
                    disp = true;
            if (disp != (pass == 1)) continue;
            if (ncls > 255) return fail(OMX_ERR_RANGE, "more than 255 distinct per-cell medium combinations");
            lut[code] = (uint8_t)ncls++;
            for (int q = 0; q < 3; ++q) {
                cls.push_back((int8_t)mm[q]);
                epsc.push_back(mm[q] ? c->media[mm[q]].eps_r : 1.0);
            }
        }
    }
    if (c->disp_tag0 >= ncls) c->disp_tag0 = 256;
    c->n_classes = ncls;
    rc = dalloc(c, (void**)&d_lut, ncode);
    if (rc) return rc;
    rc = dalloc(c, (void**)&c->d_cls, cls.size());
    if (rc) return rc;
    rc = dalloc(c, (void**)&c->d_eps_class, sizeof(double) * epsc.size());
    if (rc) return rc;
    rc = dalloc(c, (void**)&c->d_tag, L.nalloc);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(d_lut, lut.data(), ncode, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->d_cls, cls.data(), cls.size(), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->d_eps_class, epsc.data(), sizeof(double) * epsc.size(), hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_media_codes, dim3(blocks_for(ncell, 256)), dim3(256), 0, c->s, c->d_aux, L.nx, L.ny, L.nz, nm,
                       (unsigned*)nullptr, d_lut, c->d_tag, L, (int32_t*)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    c->device_bytes -= (long)(sizeof(unsigned) * ((ncode + 31) / 32) + ncode);
    dfreet(d_present);
    dfreet(d_lut);
    // the cells of dispersive classes, in memory order, for the dispersive
    // part of the E launch: i | j << 16 | k << 32 | media << 48
    dfreet(c->d_disp);
    c->n_disp = 0;
    if (c->disp_tag0 < 256) {
        if (L.nx > 65535 || L.ny > 65535 || L.nz > 65535)
            return fail(OMX_ERR_RANGE, "Drude / Drude-Lorentz media need at most 65535 cells per axis");
        std::vector<uint8_t> tags(L.nalloc);
        HIPCHK(hipMemcpyAsync(tags.data(), c->d_tag, L.nalloc, hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        std::vector<long> cells;
        for (int i = 1; i <= L.nx; ++i)
            for (int j = 1; j <= L.ny; ++j)
                for (int k = 1; k <= L.nz; ++k) {
                    const long o = L.at(i, j, k);
                    const int tg = tags[o];
                    if (tg >= c->disp_tag0) {
                        const long m0 = cls[3 * tg], m1 = cls[3 * tg + 1], m2 = cls[3 * tg + 2];
                        cells.push_back((long)i | ((long)j << 16) | ((long)k << 32) | (m0 << 48) | (m1 << 52) |
                                        (m2 << 56));
                    }
                }
        if (cells.size() > (size_t)INT32_MAX) return fail(OMX_ERR_RANGE, "too many dispersive cells");
        if (!cells.empty()) {
            rc = dalloc(c, (void**)&c->d_disp, sizeof(long) * cells.size());
            if (rc) return rc;
            HIPCHK(hipMemcpyAsync(c->d_disp, cells.data(), sizeof(long) * cells.size(), hipMemcpyHostToDevice, c->s));
            HIPCHK(hipStreamSynchronize(c->s));
            c->n_disp = (int)cells.size();
        }
    }
    return OMX_OK;
}

int media_common(omx3d_ctx* c, int n_media, const omx_medium* table) {
    if (c->media_set) return fail(OMX_ERR_STATE, "media already set for this context");
    if (n_media < 0 || n_media > OMX_MAX_MEDIA) return fail(OMX_ERR_RANGE, "n_media must be in 0..15");
    if (n_media > 0 && !table) return fail(OMX_ERR_ARG, "null media table");
    c->n_media = n_media;
    std::memset(c->media, 0, sizeof(c->media));
    c->maxp = 0;
    c->has_drude = c->has_dl = false;
    for (int m = 1; m <= n_media; ++m) {
        c->media[m] = table[m - 1];
        const omx_medium& M = c->media[m];
        if (M.kind < OMX_MEDIUM_DIELECTRIC || M.kind > OMX_MEDIUM_DL)
            return fail(OMX_ERR_ARG, "medium " + std::to_string(m) + ": unknown kind");
        if (M.kind == OMX_MEDIUM_DL && (M.n_poles < 1 || M.n_poles > OMX_MAX_POLES))
            return fail(OMX_ERR_RANGE, "medium " + std::to_string(m) + ": n_poles must be in 1..8");
        if (!(M.eps_r > 0.0)) return fail(OMX_ERR_ARG, "medium " + std::to_string(m) + ": eps_r must be > 0");
        if (M.kind == OMX_MEDIUM_DRUDE) c->has_drude = true;
        if (M.kind == OMX_MEDIUM_DL) { c->has_dl = true; c->maxp = std::max(c->maxp, (int)M.n_poles); }
    }
    int rc = dalloc(c, (void**)&c->d_med_in, sizeof(omx_medium) * (n_media + 1));
    if (rc) return rc;
    rc = dalloc(c, (void**)&c->d_med_co, sizeof(omx_medium_coeffs) * (n_media + 1));
    if (rc) return rc;
    const size_t mds = c->prec == 4 ? sizeof(MedDev<float>) : sizeof(MedDev<double>);
    rc = dalloc(c, (void**)&c->d_med, mds * (n_media + 1));
    if (rc) return rc;
    rc = dalloc(c, (void**)&c->d_aux, (size_t)c->L.nx * c->L.ny * c->L.nz);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->d_med_in, c->media, sizeof(omx_medium) * (n_media + 1), hipMemcpyHostToDevice, c->s));
    const double dt = c->d.dt, dt_eps0 = dt / kEps0;
    if (c->prec == 4)
        hipLaunchKernelGGL(k_medium_coeffs<float>, dim3(1), dim3(64), 0, c->s, c->d_med_in, n_media, dt, dt_eps0,
                           c->d_med_co, (MedDev<float>*)c->d_med);
    else
        hipLaunchKernelGGL(k_medium_coeffs<double>, dim3(1), dim3(64), 0, c->s, c->d_med_in, n_media, dt, dt_eps0,
                           c->d_med_co, (MedDev<double>*)c->d_med);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

int media_finish(omx3d_ctx* c) {
    int rc = build_tags(c);
    if (rc) return rc;
    // the dispersive state is kept for the listed cells only, in list order
    if (c->has_drude || c->has_dl) {
        rc = dalloc(c, &c->disp_state, disp_state_bytes(c));
        if (rc) return rc;
    }
    // medium kinds for the J gate of k_j_correct
    std::vector<int> kinds(OMX_MAX_MEDIA + 1, 0);
    for (int m = 1; m <= c->n_media; ++m) kinds[m] = c->media[m].kind;
    rc = dalloc(c, (void**)&c->d_kinds, sizeof(int) * kinds.size());
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->d_kinds, kinds.data(), sizeof(int) * kinds.size(), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->media_set = true;
    c->src_in_dispersive = -1;
    c->tf_fold_e = -1;
    c->wg_med = -1;
    c->patches_built = false;
    return OMX_OK;
}

}  // namespace


// ---------------------------------------------------------------------------
// Batch of independent grids: one launch per phase covers all of them.
// ---------------------------------------------------------------------------
struct omxb_ctx {
    int n = 0, prec = 8;
    std::vector<omx3d_ctx*> g;
    hipStream_t s = nullptr;
    void* d_args = nullptr;
    int* d_tile_start = nullptr;
    int total_tiles = 0;
    std::vector<int> pass_first, pass_count;
    void** d_src_field = nullptr;
    long* d_src_off = nullptr;
    double* d_src_ka = nullptr;
    double* d_src_kb = nullptr;
    int* d_src_of = nullptr;
    SourceParams* d_sp = nullptr;
    int n_src_pts = 0;
    BatchRecord* d_br = nullptr;
    int* d_grid_of = nullptr;
    int* d_local = nullptr;
    int* d_code = nullptr;
    long* d_off = nullptr;
    long n_samp = 0;
    int* d_overflow = nullptr;
    long* d_step = nullptr;
    bool periodic = false;
    int rows = 1;                   // workgroup height: the largest rows_y of the grids
    bool b32 = true;                // every grid addressed by 32-bit byte offsets
    long* d_disp = nullptr;         // cells of dispersive classes of every grid (packed as per grid)
    int* d_disp_grid = nullptr;     // grid of each workgroup of dispersive cells
    int n_disp = 0;                 // list entries (whole workgroups, padding -1)
    int disp_blocks = 0;
    std::vector<std::vector<long>> sig;   // per grid: its state when the argument blocks were built
    PatchBox* d_pbox = nullptr;     // near-field patch boxes of every grid (PatchBox::g: the grid)
    int n_pbox = 0;
    long n_pcell = 0;
    double** d_src_sh = nullptr;    // per source point: its shadow element or null
};

namespace {

template <typename P>
int bupload(hipStream_t s, P** d, const std::vector<P>& h) {
    dfreet(*d);
    if (hipMalloc((void**)d, sizeof(P) * std::max<size_t>(1, h.size())) != hipSuccess)
        return fail(OMX_ERR_NOMEM, "batch allocation failed");
    if (!h.empty()) HIPCHK(hipMemcpyAsync(*d, h.data(), sizeof(P) * h.size(), hipMemcpyHostToDevice, s));
    return OMX_OK;
}

// What the batch's argument blocks, source and detector lists are built from:
// the current field / psi buffers (the fused step swaps them), tiling,
// addressing, media, sources, detectors and the record buffer.
std::vector<long> batch_signature(const omx3d_ctx* c) {
    std::vector<long> v;
    for (int q = 0; q < 6; ++q) { v.push_back((long)(uintptr_t)c->f[q]); v.push_back((long)(uintptr_t)c->psi_h[q]); }
    v.push_back((long)(uintptr_t)c->disp_state);
    v.push_back((long)(uintptr_t)c->d_rec);
    v.push_back((long)(uintptr_t)c->d_disp);
    v.push_back(c->rows_y); v.push_back(c->xchunk); v.push_back(c->tile_order); v.push_back(c->offset_bits);
    v.push_back((long)c->sources.size()); v.push_back(c->n_samples); v.push_back(c->n_disp);
    v.push_back(c->rec_cap); v.push_back(c->media_set ? 1 : 0);
    v.push_back((long)(uintptr_t)c->d_patch); v.push_back(c->patch_cells);
    return v;
}

template <typename T>
int batch_setup(omxb_ctx* b) {
    for (auto* c : b->g) {
        int rc = ensure_patches<T>(c);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(c->s));
    }
    b->sig.clear();
    for (auto* c : b->g) b->sig.push_back(batch_signature(c));
    std::vector<YeeArgs<T>> args;
    std::vector<int> ts;
    int tot = 0;
    // each grid keeps its own tiling (rows, planes, order); its block range
    // is padded to a multiple of 8 (batch_tile)
    for (auto* c : b->g) {
        YeeArgs<T> a = make_args<T>(c);
        args.push_back(a);
        ts.push_back(tot);
        tot += ((int)yee_blocks(c, a) + 7) / 8 * 8;
    }
    ts.push_back(tot);
    b->total_tiles = tot;
    b->rows = 1;
    for (auto* c : b->g) b->rows = std::max(b->rows, c->rows_y);
    b->b32 = true;
    for (auto* c : b->g) b->b32 = b->b32 && offsets32(c);
    // cells of dispersive classes of all grids, for the first workgroups of
    // the batched E launch: each grid's cells fill whole workgroups (padded
    // with -1), so that a workgroup's grid, and with it its argument block,
    // is uniform; disp0 locates a grid's cells (and so its state) in the list
    int rc = OMX_OK;
    {
        const size_t wg = (size_t)kLanes * b->rows;
        std::vector<long> cells;
        std::vector<int> gob;
        for (int q = 0; q < b->n; ++q) {
            omx3d_ctx* c = b->g[q];
            if (!(c->media_set && c->n_classes > 1 && c->n_disp > 0)) continue;
            std::vector<long> h(c->n_disp);
            HIPCHK(hipMemcpy(h.data(), c->d_disp, sizeof(long) * h.size(), hipMemcpyDeviceToHost));
            args[q].disp0 = (long)cells.size();
            cells.insert(cells.end(), h.begin(), h.end());
            cells.resize((cells.size() + wg - 1) / wg * wg, -1L);
            gob.resize(cells.size() / wg, q);
        }
        if (cells.size() > (size_t)INT32_MAX) return fail(OMX_ERR_RANGE, "too many dispersive cells in the batch");
        b->n_disp = (int)cells.size();
        b->disp_blocks = (int)gob.size();
        rc = bupload(b->s, &b->d_disp, cells);
        if (!rc) rc = bupload(b->s, &b->d_disp_grid, gob);
        if (rc) return rc;
    }
    rc = bupload(b->s, (YeeArgs<T>**)&b->d_args, args);
    if (!rc) rc = bupload(b->s, &b->d_tile_start, ts);
    if (rc) return rc;
    // near-field patch boxes of all grids, one flat index over them
    {
        std::vector<PatchBox> pb;
        long first = 0;
        for (int q = 0; q < b->n; ++q)
            for (PatchBox B : b->g[q]->h_boxes) {
                B.g = q;
                B.first = first;
                first += patch_padded((long)B.w[0] * B.w[1] * B.w[2]);
                pb.push_back(B);
            }
        b->n_pbox = (int)pb.size();
        b->n_pcell = first;
        rc = bupload(b->s, &b->d_pbox, pb);
        if (rc) return rc;
    }
    // sources: one pass per repeated (field, cell), in source order
    std::map<std::pair<void*, long>, int> seen;
    std::vector<std::vector<void*>> pf;
    std::vector<std::vector<long>> po;
    std::vector<std::vector<double>> pa, pb;
    std::vector<std::vector<int>> ps;
    std::vector<std::vector<double*>> psh;
    std::vector<SourceParams> sp;
    for (auto* c : b->g) {
        for (auto& src : c->sources) {
            const int sid = (int)sp.size();
            sp.push_back(source_params(c, src));
            std::vector<long> off(src.npts);
            std::vector<double> ka(src.npts), kb(src.npts);
            HIPCHK(hipMemcpy(off.data(), src.d_off, sizeof(long) * src.npts, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(ka.data(), src.d_ka, sizeof(double) * src.npts, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(kb.data(), src.d_kb, sizeof(double) * src.npts, hipMemcpyDeviceToHost));
            void* fp = c->f[src.dir - 1];
            for (int q = 0; q < src.npts; ++q) {
                const int pass = seen[std::make_pair(fp, off[q])]++;
                if ((int)pf.size() <= pass) {
                    pf.resize(pass + 1); po.resize(pass + 1); pa.resize(pass + 1); pb.resize(pass + 1);
                    ps.resize(pass + 1); psh.resize(pass + 1);
                }
                pf[pass].push_back(fp); po[pass].push_back(off[q]); pa[pass].push_back(ka[q]);
                pb[pass].push_back(kb[q]); ps[pass].push_back(sid); psh[pass].push_back(src.h_sh[q]);
            }
        }
    }
    std::vector<void*> F; std::vector<long> O; std::vector<double> A, B; std::vector<int> S;
    std::vector<double*> SH;
    b->pass_first.clear(); b->pass_count.clear();
    for (size_t q = 0; q < pf.size(); ++q) {
        b->pass_first.push_back((int)F.size());
        b->pass_count.push_back((int)pf[q].size());
        F.insert(F.end(), pf[q].begin(), pf[q].end()); O.insert(O.end(), po[q].begin(), po[q].end());
        A.insert(A.end(), pa[q].begin(), pa[q].end()); B.insert(B.end(), pb[q].begin(), pb[q].end());
        S.insert(S.end(), ps[q].begin(), ps[q].end()); SH.insert(SH.end(), psh[q].begin(), psh[q].end());
    }
    b->n_src_pts = (int)F.size();
    rc = bupload(b->s, &b->d_src_field, F);
    if (!rc) rc = bupload(b->s, &b->d_src_off, O);
    if (!rc) rc = bupload(b->s, &b->d_src_ka, A);
    if (!rc) rc = bupload(b->s, &b->d_src_kb, B);
    if (!rc) rc = bupload(b->s, &b->d_src_of, S);
    if (!rc) rc = bupload(b->s, &b->d_src_sh, SH);
    if (!rc) rc = bupload(b->s, &b->d_sp, sp);
    if (rc) return rc;
    // detectors
    std::vector<int> go, lo, cd;
    std::vector<long> of;
    for (int q = 0; q < b->n; ++q) {
        omx3d_ctx* c = b->g[q];
        for (long k = 0; k < c->n_samples; ++k) {
            go.push_back(q); lo.push_back((int)k); cd.push_back(c->h_det_field[k]); of.push_back(c->h_det_off[k]);
        }
    }
    b->n_samp = (long)go.size();
    rc = bupload(b->s, &b->d_grid_of, go);
    if (!rc) rc = bupload(b->s, &b->d_local, lo);
    if (!rc) rc = bupload(b->s, &b->d_code, cd);
    if (!rc) rc = bupload(b->s, &b->d_off, of);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->s));
    return OMX_OK;
}

template <typename T>
int batch_step_launches(omxb_ctx* b) {
    const dim3 blk(kLanes, b->rows);
    if (b->periodic)
        for (auto* c : b->g) { int rc = wrap_e<T>(c, b->s); if (rc) return rc; }
    if (b->b32)
        hipLaunchKernelGGL((k_h_update_batched<T, true>), dim3(b->total_tiles), blk, 0, b->s,
                           (const YeeArgs<T>*)b->d_args, b->d_tile_start, b->n);
    else
        hipLaunchKernelGGL((k_h_update_batched<T, false>), dim3(b->total_tiles), blk, 0, b->s,
                           (const YeeArgs<T>*)b->d_args, b->d_tile_start, b->n);
    if (b->n_pcell)
        hipLaunchKernelGGL(k_patch_batched<T>, dim3(blocks_for(b->n_pcell, 256)), dim3(256), 0, b->s,
                           (const YeeArgs<T>*)b->d_args, b->d_pbox, b->n_pbox, b->n_pcell, 0);
    for (size_t q = 0; q < b->pass_first.size(); ++q) {
        const int f0 = b->pass_first[q], nq = b->pass_count[q];
        hipLaunchKernelGGL(k_source_batched<T>, dim3(blocks_for(nq, 256)), dim3(256), 0, b->s,
                           (T* const*)(b->d_src_field + f0), b->d_src_off + f0, b->d_src_ka + f0, b->d_src_kb + f0,
                           b->d_src_of + f0, b->d_sp, nq, (const long*)b->d_step,
                           (double* const*)(b->d_src_sh + f0));
    }
    if (b->periodic)
        for (auto* c : b->g) { int rc = wrap_h<T>(c, b->s); if (rc) return rc; }
    // the dispersive cells of every grid ride in the first db workgroups
    const int db = b->disp_blocks;
    if (b->b32)
        hipLaunchKernelGGL((k_e_update_batched<T, true>), dim3(b->total_tiles + db), blk, 0, b->s,
                           (const YeeArgs<T>*)b->d_args, b->d_tile_start, b->n, b->d_disp, b->d_disp_grid,
                           b->n_disp, db);
    else
        hipLaunchKernelGGL((k_e_update_batched<T, false>), dim3(b->total_tiles + db), blk, 0, b->s,
                           (const YeeArgs<T>*)b->d_args, b->d_tile_start, b->n, b->d_disp, b->d_disp_grid,
                           b->n_disp, db);
    if (b->n_pcell)
        hipLaunchKernelGGL(k_patch_batched<T>, dim3(blocks_for(b->n_pcell, 256)), dim3(256), 0, b->s,
                           (const YeeArgs<T>*)b->d_args, b->d_pbox, b->n_pbox, b->n_pcell, 1);
    if (b->n_samp > 0)
        hipLaunchKernelGGL(k_sample_batched<T>, dim3(blocks_for(b->n_samp, 256)), dim3(256), 0, b->s,
                           (const YeeArgs<T>*)b->d_args, b->d_br, b->d_grid_of, b->d_local, b->d_code, b->d_off,
                           b->n_samp, (const long*)b->d_step, b->d_overflow);
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, b->s, b->d_step);
    HIPCHK(hipGetLastError());
    return OMX_OK;
}

template <typename T>
int batch_run(omxb_ctx* b, int tt_first, int n_steps, int every, int use_graph) {
    // the grids' state may have changed since the argument blocks were built
    // (a fused single-grid run swaps the field buffers; sources, detectors or
    // the record buffer may have been added): build them again then
    for (int q = 0; q < b->n; ++q) {
        omx3d_ctx* c = b->g[q];
        int rc = upload_detectors(c);
        if (!rc) rc = ensure_patches<T>(c);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(c->s));
    }
    bool stale = false;
    for (int q = 0; q < b->n; ++q) stale = stale || batch_signature(b->g[q]) != b->sig[q];
    if (stale) {
        int rc = batch_setup<T>(b);
        if (rc) return rc;
    }
    std::vector<BatchRecord> br(b->n);
    bool found = false;
    int first_tt = 0;
    if (every > 0)
        for (int tt = tt_first; tt < tt_first + n_steps; ++tt)
            if (tt % every == 0) { first_tt = tt; found = true; break; }
    int nrec = 0;
    if (every > 0)
        for (int tt = tt_first; tt < tt_first + n_steps; ++tt) nrec += (tt % every == 0);
    for (int q = 0; q < b->n; ++q) {
        omx3d_ctx* c = b->g[q];
        br[q].rec = c->d_rec;
        br[q].n_samples = c->n_samples;
        br[q].every = c->n_samples > 0 ? every : 0;
        // row of step tt: tt/every - row0, so that the first recorded step
        // of this call lands on the grid's next free row
        br[q].row0 = found ? first_tt / every - c->rec_rows : 0;
        br[q].cap = c->rec_cap;
        // as omx3d_run: recording into a full (or absent) buffer is an error
        if (br[q].every > 0 && nrec > 0 && c->rec_rows + nrec > c->rec_cap)
            return fail(OMX_ERR_RANGE, "batch grid " + std::to_string(q) + ": detector record buffer full");
    }
    int rc = bupload(b->s, &b->d_br, br);
    if (rc) return rc;
    const long st = tt_first;
    HIPCHK(hipMemcpyAsync(b->d_step, &st, sizeof(long), hipMemcpyHostToDevice, b->s));
    const int zero = 0;
    HIPCHK(hipMemcpyAsync(b->d_overflow, &zero, sizeof(int), hipMemcpyHostToDevice, b->s));
    if (use_graph && n_steps > 0) {
        hipGraph_t graph;
        hipGraphExec_t exec;
        HIPCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
        rc = batch_step_launches<T>(b);
        hipGraph_t gtmp;
        const hipError_t ec = hipStreamEndCapture(b->s, &gtmp);
        if (rc) return rc;
        if (ec != hipSuccess) return fail(OMX_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
        graph = gtmp;
        HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        for (int t = 0; t < n_steps; ++t) HIPCHK(hipGraphLaunch(exec, b->s));
        HIPCHK(hipStreamSynchronize(b->s));
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
    } else {
        for (int t = 0; t < n_steps; ++t) {
            rc = batch_step_launches<T>(b);
            if (rc) return rc;
        }
        HIPCHK(hipStreamSynchronize(b->s));
    }
    int ov = 0;
    HIPCHK(hipMemcpy(&ov, b->d_overflow, sizeof(int), hipMemcpyDeviceToHost));
    if (ov) return fail(OMX_ERR_RANGE, "batch detector rows out of range");
    for (auto* c : b->g)
        if (c->n_samples > 0) c->rec_rows += nrec;
    return OMX_OK;
}

}  // namespace
// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* omx_last_error(void) { return g_err.c_str(); }

int omx_version(void) { return 1; }

// get_DL_library_data (classical_medium_mod.f90:125-199).  The reference's
// literals are default REAL: products of two literals are single precision,
// a literal times a double is the single value promoted.
int omx_medium_library(const char* name, double eps_r, omx_medium* out) {
    if (!name || !out) return fail(OMX_ERR_ARG, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->kind = OMX_MEDIUM_DL;
    out->eps_r = eps_r;
    const double ev = kEvToAu;
    auto F = [](float v) { return (double)v; };
    double wp;
    std::string n(name);
    if (n == "Ag") {
        out->n_poles = 5;
        out->gamma_d = F(0.048f) * ev; wp = F(9.01f) * ev;
        out->omega_d = F(9.01f * 0.919238815542512f) * ev;
        const float w[5] = {0.816f, 4.481f, 8.185f, 9.083f, 20.29f};
        const float g[5] = {3.886f, 0.452f, 0.065f, 0.916f, 2.419f};
        const float f[5] = {0.065f, 0.124f, 0.011f, 0.840f, 5.646f};
        for (int q = 0; q < 5; ++q) { out->omega_k[q] = F(w[q]) * ev; out->gamma_k[q] = F(g[q]) * ev; out->a_k[q] = F(f[q]) * (wp * wp); }
    } else if (n == "Al") {
        out->n_poles = 4;
        out->gamma_d = F(0.047f) * ev; wp = F(14.98f) * ev;
        out->omega_d = F(14.98f * 0.723187389270582f) * ev;
        const float w[4] = {0.162f, 1.544f, 1.808f, 3.473f};
        const float g[4] = {0.333f, 0.312f, 1.351f, 3.382f};
        const float f[4] = {0.227f, 0.050f, 0.166f, 0.030f};
        for (int q = 0; q < 4; ++q) { out->omega_k[q] = F(w[q]) * ev; out->gamma_k[q] = F(g[q]) * ev; out->a_k[q] = F(f[q]) * (wp * wp); }
    } else if (n == "Au") {
        out->n_poles = 5;
        out->gamma_d = F(0.053f) * ev; wp = F(9.03f) * ev;
        out->omega_d = F(9.03f * 0.871779788708135f) * ev;
        const float w[5] = {0.415f, 0.830f, 2.969f, 4.304f, 13.32f};
        const float g[5] = {0.241f, 0.345f, 0.870f, 2.494f, 2.214f};
        const float f[5] = {0.024f, 0.010f, 0.071f, 0.601f, 4.384f};
        for (int q = 0; q < 5; ++q) { out->omega_k[q] = F(w[q]) * ev; out->gamma_k[q] = F(g[q]) * ev; out->a_k[q] = F(f[q]) * (wp * wp); }
    } else {
        return fail(OMX_ERR_ARG, "unknown Drude-Lorentz library material '" + n + "' (Ag, Al, Au)");
    }
    return OMX_OK;
}

int omx3d_create(const omx3d_desc* desc, omx3d_ctx** out) {
    if (!desc || !out) return fail(OMX_ERR_ARG, "null argument");
    *out = nullptr;
    const omx3d_desc& d = *desc;
    if (d.nx < 2 || d.ny < 2 || d.nz < 2) return fail(OMX_ERR_ARG, "grid must be at least 2 cells per axis");
    if (d.precision != OMX_FP32 && d.precision != OMX_FP64) return fail(OMX_ERR_ARG, "precision must be 4 or 8");
    if (!(d.dt > 0.0) || !(d.dr > 0.0)) return fail(OMX_ERR_ARG, "dt and dr must be positive");
    unsigned mask = 0;
    const int n[3] = {d.nx, d.ny, d.nz};
    for (int a = 0; a < 3; ++a) {
        const int b = d.boundaries[a];
        if (b != OMX_BC_CLOSE && b != OMX_BC_PERIODIC && b != OMX_BC_CPML)
            return fail(OMX_ERR_ARG, "boundary code of axis " + std::to_string(a + 1) + " must be 1, 2 or 3");
        if (b == OMX_BC_CPML) {
            if (d.npml < 2) return fail(OMX_ERR_ARG, "CPML needs npml >= 2");
            if (2 * d.npml > n[a]) return fail(OMX_ERR_ARG, "CPML thicker than half the grid on axis " + std::to_string(a + 1));
            mask |= 3u << (2 * a);
        }
    }
    omx3d_ctx* c = new omx3d_ctx();
    c->d = d;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return fail(OMX_ERR_HIP, "hipGetDevice failed");
    }
    c->prec = d.precision;
    c->esz = (size_t)d.precision;
    // default tile shapes, measured fastest on 256^3 CPML grids: FP32 4 rows
    // x 4 planes, FP64 2 rows x 4 planes, whole row blocks per XCD
    c->rows_y = d.precision == OMX_FP32 ? 4 : 2;
    c->xchunk = 4;
    c->tile_order = 2;
    // fused step: FP32 4 rows x 16 planes per wavefront, FP64 2 x 8
    c->wg_rows = d.precision == OMX_FP32 ? 2 : 1;
    c->cpml_mask = mask;
    Layout& L = c->L;
    const int A = 128 / (int)c->esz;
    L.nx = d.nx; L.ny = d.ny; L.nz = d.nz;
    L.sy = d.ny + 2;
    L.sz = ((d.nz + 2 + A - 1) / A) * A;
    L.sx = (long)L.sy * L.sz;
    L.off = A - 1;
    L.nalloc = (long)(d.nx + 2) * L.sx + A;
    int rc = OMX_OK;
    auto bail = [&](int r) { omx3d_destroy(c); return r; };
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(OMX_ERR_HIP, "hipStreamCreate failed"));
    for (int q = 0; q < 6; ++q)
        if ((rc = dalloc(c, &c->f[q], c->esz * L.nalloc))) return bail(rc);

    // coefficient tables
    const int np = d.npml > 0 ? d.npml : 1;
    c->co_be = 0; c->co_ce = np + 1; c->co_bh = 2 * (np + 1); c->co_ch = c->co_bh + np;
    long p = c->co_ch + np;
    for (int q = 0; q < 6; ++q) { c->co_den[q] = p; p += n[q % 3] + 2; }
    for (int q = 0; q < 6; ++q) { c->co_cd[q] = p; p += n[q % 3] + 2; }
    c->ncoef = p;
    if (np > 64 && mask) return bail(fail(OMX_ERR_RANGE, "npml above 64 is not supported"));
    if ((rc = dalloc(c, (void**)&c->dcoef, sizeof(double) * c->ncoef))) return bail(rc);
    if ((rc = dalloc(c, &c->tcoef, c->esz * c->ncoef))) return bail(rc);
    CpmlProfile pr;
    pr.npml = d.npml; pr.dt = d.dt; pr.dr = d.dr;
    const double sig_to_au = ((kCToAu * kCToAu) * kSecToAu) / (kKgToAu * ((kMToAu * kMToAu) * kMToAu));
    pr.sig_max = (double)(0.8f * (float)(kGradeM + 1)) / (d.dr * std::sqrt(kMu0 / kEps0));
    pr.alpha_scale = kAlphaCPML * sig_to_au;
    const int kmax = std::max(std::max(d.nx, d.ny), std::max(d.nz, np));
    double* dc = c->dcoef;
    hipLaunchKernelGGL(k_cpml_coeffs, dim3(blocks_for(kmax, 128)), dim3(128), 0, c->s, pr, d.nx, d.ny, d.nz, mask,
                       dc + c->co_be, dc + c->co_ce, dc + c->co_bh, dc + c->co_ch, dc + c->co_den[0], dc + c->co_den[1],
                       dc + c->co_den[2], dc + c->co_den[3], dc + c->co_den[4], dc + c->co_den[5]);
    if (hipGetLastError() != hipSuccess) return bail(fail(OMX_ERR_HIP, "coefficient kernel launch failed"));
    std::vector<double> hc(c->ncoef);
    if (hipMemcpyAsync(hc.data(), dc, sizeof(double) * c->ncoef, hipMemcpyDeviceToHost, c->s) != hipSuccess ||
        hipStreamSynchronize(c->s) != hipSuccess)
        return bail(fail(OMX_ERR_HIP, "coefficient readback failed"));
    for (long q = c->co_ce; q < c->co_ce + np + 1; ++q) hc[q] /= d.dr;  // ce/dr
    for (long q = c->co_ch; q < c->co_ch + np; ++q) hc[q] /= d.dr;      // ch/dr
    for (int q = 0; q < 6; ++q)                                           // dt_eps0*den_e, dt_mu0*den_h
        for (int k = 0; k < n[q % 3] + 2; ++k)
            hc[c->co_cd[q] + k] = (q < 3 ? d.dt / kEps0 : d.dt / kMu0) * hc[c->co_den[q] + k];
    if (c->prec == 4) {
        std::vector<float> hf(c->ncoef);
        for (long q = 0; q < c->ncoef; ++q) hf[q] = (float)hc[q];
        if (hipMemcpyAsync(c->tcoef, hf.data(), sizeof(float) * c->ncoef, hipMemcpyHostToDevice, c->s) != hipSuccess ||
            hipStreamSynchronize(c->s) != hipSuccess)
            return bail(fail(OMX_ERR_HIP, "coefficient upload failed"));
        float* tf = (float*)c->tcoef;
        hipLaunchKernelGGL(k_cpml_bm1, dim3(1), dim3(128), 0, c->s, pr, mask, tf + c->co_be, tf + c->co_bh);
    } else {
        if (hipMemcpyAsync(c->tcoef, hc.data(), sizeof(double) * c->ncoef, hipMemcpyHostToDevice, c->s) != hipSuccess)
            return bail(fail(OMX_ERR_HIP, "coefficient upload failed"));
    }

    // CPML psi arrays (mxll_3D_mod.f90:236-276), non-MPI: two sections per axis.
    if (mask) {
        c->nsh = 2 * (d.npml - 1);
        c->nse = 2 * d.npml;
        const size_t A2 = (size_t)A;
        const size_t xs[2] = {c->nsh * L.sx + A2, c->nse * L.sx + A2};
        const size_t ys[2] = {(size_t)(d.nx + 2) * c->nsh * L.sz + A2, (size_t)(d.nx + 2) * c->nse * L.sz + A2};
        const size_t zs[2] = {(size_t)zpsi_elems(L.sy, d.nx, d.npml), (size_t)zpsi_elems(L.sy, d.nx, d.npml)};
        for (int a = 0; a < 3; ++a) {
            if (!((mask >> (2 * a)) & 3u)) continue;
            const size_t* sz = a == 0 ? xs : (a == 1 ? ys : zs);
            for (int q = 0; q < 2; ++q) {
                if ((rc = dalloc(c, &c->psi_h[2 * a + q], c->esz * sz[0]))) return bail(rc);
                if ((rc = dalloc(c, &c->psi_e[2 * a + q], c->esz * sz[1]))) return bail(rc);
            }
        }
    }
    if (hipStreamSynchronize(c->s) != hipSuccess) return bail(fail(OMX_ERR_HIP, "create: stream sync failed"));
    *out = c;
    return OMX_OK;
}

int omx3d_destroy(omx3d_ctx* c) {
    if (!c) return OMX_OK;
    if (c->s) (void)hipStreamSynchronize(c->s);
    free_patches(c);
    for (auto& p : c->f) dfree(p);
    for (auto& p : c->f_alt) dfree(p);
    for (auto& p : c->psi_h_alt) dfree(p);
    dfreet(c->dcoef); dfree(c->tcoef);
    for (auto& p : c->psi_h) dfree(p);
    for (auto& p : c->psi_e) dfree(p);
    dfreet(c->d_aux); dfreet(c->d_tag); dfreet(c->d_cls); dfreet(c->d_disp); dfree(c->d_med); dfreet(c->d_med_in); dfreet(c->d_med_co);
    dfreet(c->d_eps_class); dfreet(c->d_kinds); dfree(c->disp_state);
    for (int q = 0; q < 3; ++q) {
        dfreet(c->d_aff_off[q]); dfreet(c->d_aff_a[q]); dfreet(c->d_aff_b[q]);
    }
    for (auto& s : c->sources) { dfreet(s.d_off); dfreet(s.d_ka); dfreet(s.d_kb); }
    dfreet(c->d_det_field); dfreet(c->d_det_off); dfreet(c->d_rec);
    dfreet(c->d_q_off); dfreet(c->d_q_slot); dfreet(c->d_J); dfreet(c->d_q_e); dfreet(c->d_q_dpdt); dfreet(c->d_tls);
    dfreet(c->d_tf); dfreet(c->d_red);
    for (auto e : c->ev) (void)hipEventDestroy(e);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
    return OMX_OK;
}

int omx3d_set_tiling(omx3d_ctx* c, int32_t rows_y, int32_t xchunk, int32_t tile_order) {
    CHECK_CTX(c);
    if (tile_order < -1 || tile_order > 2) return fail(OMX_ERR_ARG, "tile_order must be -1, 0, 1 or 2");
    if (rows_y != 0) {
        if (rows_y != 1 && rows_y != 2 && rows_y != 4) return fail(OMX_ERR_ARG, "rows_y must be 1, 2 or 4");
        c->rows_y = rows_y;
    }
    if (xchunk != 0) {
        if (xchunk < 1) return fail(OMX_ERR_ARG, "xchunk must be >= 1");
        c->xchunk = xchunk;
    }
    if (tile_order >= 0) c->tile_order = tile_order;
    return OMX_OK;
}

int omx3d_set_addressing(omx3d_ctx* c, int32_t offset_bits) {
    CHECK_CTX(c);
    if (offset_bits != 32 && offset_bits != 64) return fail(OMX_ERR_ARG, "offset_bits must be 32 or 64");
    c->offset_bits = offset_bits;
    return OMX_OK;
}

int omx3d_get_tiling(omx3d_ctx* c, int32_t* rows_y, int32_t* xchunk, int32_t* tile_order) {
    CHECK_CTX(c);
    if (rows_y) *rows_y = c->rows_y;
    if (xchunk) *xchunk = c->xchunk;
    if (tile_order) *tile_order = c->tile_order;
    return OMX_OK;
}

int omx3d_get_cpml(omx3d_ctx* c, double* be, double* ce, double* bh, double* ch, double* den) {
    CHECK_CTX(c);
    std::vector<double> h(c->ncoef);
    HIPCHK(hipMemcpyAsync(h.data(), c->dcoef, sizeof(double) * c->ncoef, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    const int np = c->d.npml;
    if (c->cpml_mask) {
        for (int k = 1; k <= np; ++k) {
            if (be) be[k - 1] = h[c->co_be + k];
            if (ce) ce[k - 1] = h[c->co_ce + k];
        }
        for (int k = 1; k <= np - 1; ++k) {
            if (bh) bh[k - 1] = h[c->co_bh + k];
            if (ch) ch[k - 1] = h[c->co_ch + k];
        }
    }
    if (den) {
        const int n[3] = {c->d.nx, c->d.ny, c->d.nz};
        long w = 0;
        for (int q = 0; q < 6; ++q)
            for (int k = 1; k <= n[q % 3]; ++k) den[w++] = h[c->co_den[q] + k];
    }
    return OMX_OK;
}

int omx3d_set_media(omx3d_ctx* c, int32_t n_media, const omx_medium* table, const int8_t* aux) {
    CHECK_CTX(c);
    if (!aux) return fail(OMX_ERR_ARG, "null aux map");
    const long ncell = (long)c->L.nx * c->L.ny * c->L.nz;
    for (long q = 0; q < ncell; ++q)
        if (aux[q] < 0 || aux[q] > n_media) return fail(OMX_ERR_ARG, "aux map value out of 0..n_media");
    int rc = media_common(c, n_media, table);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->d_aux, aux, ncell, hipMemcpyHostToDevice, c->s));
    return media_finish(c);
}

int omx3d_set_media_shapes(omx3d_ctx* c, int32_t n_media, const omx_medium* table, int32_t n_shapes,
                           const omx_shape* shapes) {
    CHECK_CTX(c);
    if (n_shapes < 0 || (n_shapes > 0 && !shapes)) return fail(OMX_ERR_ARG, "bad shape list");
    for (int s = 0; s < n_shapes; ++s)
        if (shapes[s].medium < 0 || shapes[s].medium > n_media || shapes[s].type < 1 || shapes[s].type > 4)
            return fail(OMX_ERR_ARG, "shape " + std::to_string(s) + ": bad type or medium");
    int rc = media_common(c, n_media, table);
    if (rc) return rc;
    omx_shape* d_sh = nullptr;
    rc = dalloc(c, (void**)&d_sh, sizeof(omx_shape) * std::max(1, (int)n_shapes));
    if (rc) return rc;
    if (n_shapes) HIPCHK(hipMemcpyAsync(d_sh, shapes, sizeof(omx_shape) * n_shapes, hipMemcpyHostToDevice, c->s));
    const long ncell = (long)c->L.nx * c->L.ny * c->L.nz;
    hipLaunchKernelGGL(k_rasterize, dim3(blocks_for(ncell, 256)), dim3(256), 0, c->s, d_sh, n_shapes, c->L.nx, c->L.ny,
                       c->L.nz, c->d_aux);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    c->device_bytes -= (long)(sizeof(omx_shape) * std::max(1, (int)n_shapes));
    dfreet(d_sh);
    return media_finish(c);
}

int omx3d_get_media_map(omx3d_ctx* c, int32_t* map) {
    CHECK_CTX(c);
    if (!map) return fail(OMX_ERR_ARG, "null output");
    const long ncell = (long)c->L.nx * c->L.ny * c->L.nz;
    if (!c->media_set) {
        std::memset(map, 0, sizeof(int32_t) * 3 * ncell);
        return OMX_OK;
    }
    int32_t* d_map = nullptr;
    int rc = dalloc(c, (void**)&d_map, sizeof(int32_t) * 3 * ncell);
    if (rc) return rc;
    hipLaunchKernelGGL(k_media_codes, dim3(blocks_for(ncell, 256)), dim3(256), 0, c->s, c->d_aux, c->L.nx, c->L.ny,
                       c->L.nz, c->n_media, (unsigned*)nullptr, (const uint8_t*)nullptr, (uint8_t*)nullptr, c->L, d_map);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(map, d_map, sizeof(int32_t) * 3 * ncell, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->device_bytes -= (long)(sizeof(int32_t) * 3 * ncell);
    dfreet(d_map);
    return OMX_OK;
}

int omx3d_get_medium_coeffs(omx3d_ctx* c, int32_t m, omx_medium_coeffs* out) {
    CHECK_CTX(c);
    if (!c->media_set) return fail(OMX_ERR_STATE, "no media set");
    if (m < 1 || m > c->n_media || !out) return fail(OMX_ERR_ARG, "medium index out of range");
    HIPCHK(hipMemcpyAsync(out, c->d_med_co + m, sizeof(omx_medium_coeffs), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return OMX_OK;
}

int omx3d_step_h(omx3d_ctx* c) {
    CHECK_CTX(c);
    return c->prec == 4 ? launch_h<float>(c, false) : launch_h<double>(c, false);
}

int omx3d_step_e(omx3d_ctx* c, double time, double t_skip) {
    CHECK_CTX(c);
    return c->prec == 4 ? launch_e<float>(c, time, t_skip, false, 0.0) : launch_e<double>(c, time, t_skip, false, 0.0);
}

int omx3d_add_source(omx3d_ctx* c, int32_t dir, double j_amp, double w0, double t0, double tau, double t_init,
                     double t_final, double phase, int32_t n_ker, const int32_t* ind, const double* ker) {
    CHECK_CTX(c);
    if (dir < 1 || dir > 3) return fail(OMX_ERR_ARG, "source direction must be 1, 2 or 3");
    if (n_ker < 0 || !ind || !ker) return fail(OMX_ERR_ARG, "bad source kernel");
    const int w = 2 * n_ker + 1;
    const long np = (long)w * w * w;
    std::vector<long> off(np);
    std::vector<double> ka(np), kb(np);
    const int n[3] = {c->d.nx, c->d.ny, c->d.nz};
    for (int k = 0; k < w; ++k)
        for (int j = 0; j < w; ++j)
            for (int i = 0; i < w; ++i) {
                const long p = i + (long)w * (j + (long)w * k);
                const int ii = ind[p], jj = ind[p + np], kk = ind[p + 2 * np];
                const int idx[3] = {ii, jj, kk};
                for (int a = 0; a < 3; ++a)
                    if (idx[a] < 1 || idx[a] > n[a]) return fail(OMX_ERR_RANGE, "source point outside the grid");
                off[p] = c->L.at(ii, jj, kk);
                ka[p] = ker[p];
                int ni = i, nj = j, nk = k;
                if (dir == 1) ++ni; else if (dir == 2) ++nj; else ++nk;
                kb[p] = (ni < w && nj < w && nk < w) ? ker[ni + (long)w * (nj + (long)w * nk)] : 0.0;
            }
    SourceDev s;
    s.dir = dir; s.j_amp = j_amp; s.w0 = w0; s.t0 = t0; s.tau = tau; s.t_init = t_init; s.t_final = t_final;
    s.phase = phase; s.npts = (int)np;
    s.d_off = nullptr; s.d_ka = nullptr; s.d_kb = nullptr; s.d_sh = nullptr;
    s.h_ijk.resize(3 * (size_t)np);
    for (long p = 0; p < np; ++p)
        for (int a = 0; a < 3; ++a) s.h_ijk[3 * p + a] = ind[p + a * np];
    s.h_sh.assign(np, nullptr);
    int rc = dalloc(c, (void**)&s.d_off, sizeof(long) * np);
    if (!rc) rc = dalloc(c, (void**)&s.d_ka, sizeof(double) * np);
    if (!rc) rc = dalloc(c, (void**)&s.d_kb, sizeof(double) * np);
    if (rc) { dfreet(s.d_off); dfreet(s.d_ka); dfreet(s.d_kb); return rc; }
    HIPCHK(hipMemcpyAsync(s.d_off, off.data(), sizeof(long) * np, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(s.d_ka, ka.data(), sizeof(double) * np, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(s.d_kb, kb.data(), sizeof(double) * np, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->sources.push_back(s);
    c->src_in_dispersive = -1;
    c->patches_built = false;
    return OMX_OK;
}

int omx3d_inject_sources(omx3d_ctx* c, double time) {
    CHECK_CTX(c);
    return c->prec == 4 ? launch_sources<float>(c, time) : launch_sources<double>(c, time);
}

int omx3d_add_detector(omx3d_ctx* c, int32_t det_type, int32_t field, int32_t nd, const int32_t* idx) {
    CHECK_CTX(c);
    if (det_type < OMX_DET_POINT || det_type > OMX_DET_VOLUME) return fail(OMX_ERR_ARG, "bad detector type");
    if (field < OMX_EX || field > OMX_HZ) return fail(OMX_ERR_ARG, "bad detector field");
    if (nd < 0 || (nd > 0 && !idx)) return fail(OMX_ERR_ARG, "bad detector index list");
    if (c->rec_cap > 0) {  // the row width changes: the record buffer is sized again by the caller
        HIPCHK(hipStreamSynchronize(c->s));
        c->device_bytes -= (long)(sizeof(double) * (size_t)c->rec_cap * c->n_samples);
        dfreet(c->d_rec);
        c->rec_cap = c->rec_rows = 0;
    }
    const int n[3] = {c->d.nx, c->d.ny, c->d.nz};
    for (int q = 0; q < nd; ++q) {
        const int ii = idx[q], jj = idx[q + nd], kk = idx[q + 2 * nd];
        const int v[3] = {ii, jj, kk};
        for (int a = 0; a < 3; ++a)
            if (v[a] < 1 || v[a] > n[a]) return fail(OMX_ERR_RANGE, "detector index outside the grid");
        c->h_det_field.push_back(2 * field + (det_type == OMX_DET_POINT ? 1 : 0));
        c->h_det_off.push_back(c->L.at(ii, jj, kk));
    }
    DetectorHost dh;
    dh.type = det_type; dh.field = field; dh.nd = nd; dh.first = c->n_samples;
    c->dets.push_back(dh);
    c->n_samples += nd;
    c->det_dirty = true;
    return OMX_OK;
}

int omx3d_set_record_capacity(omx3d_ctx* c, int32_t rows) {
    CHECK_CTX(c);
    if (rows < 0) return fail(OMX_ERR_ARG, "rows must be >= 0");
    dfreet(c->d_rec);
    c->rec_cap = 0;
    c->rec_rows = 0;
    if (rows > 0 && c->n_samples > 0) {
        int rc = dalloc(c, (void**)&c->d_rec, sizeof(double) * (size_t)rows * c->n_samples);
        if (rc) return rc;
        c->rec_cap = rows;
    }
    return upload_detectors(c);
}

int omx3d_record(omx3d_ctx* c) {
    CHECK_CTX(c);
    int rc = upload_detectors(c);
    if (rc) return rc;
    if (c->n_samples == 0) return OMX_OK;
    if (c->rec_rows >= c->rec_cap) return fail(OMX_ERR_RANGE, "detector record buffer full");
    double* dst = c->d_rec + (long)c->rec_rows * c->n_samples;
    rc = c->prec == 4 ? launch_record<float>(c, dst) : launch_record<double>(c, dst);
    if (rc) return rc;
    c->rec_rows++;
    return OMX_OK;
}

int omx3d_read_records(omx3d_ctx* c, double* out, int32_t* rows) {
    CHECK_CTX(c);
    if (rows) *rows = c->rec_rows;
    if (out && c->rec_rows > 0)
        HIPCHK(hipMemcpyAsync(out, c->d_rec, sizeof(double) * (size_t)c->rec_rows * c->n_samples, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return OMX_OK;
}

int omx3d_sample(omx3d_ctx* c, double* out) {
    CHECK_CTX(c);
    if (!out) return fail(OMX_ERR_ARG, "null output");
    int rc = upload_detectors(c);
    if (rc) return rc;
    if (c->n_samples == 0) return OMX_OK;
    double* tmp = nullptr;
    rc = dalloc(c, (void**)&tmp, sizeof(double) * c->n_samples);
    if (rc) return rc;
    rc = c->prec == 4 ? launch_record<float>(c, tmp) : launch_record<double>(c, tmp);
    if (!rc) {
        if (hipMemcpyAsync(out, tmp, sizeof(double) * c->n_samples, hipMemcpyDeviceToHost, c->s) != hipSuccess ||
            hipStreamSynchronize(c->s) != hipSuccess)
            rc = fail(OMX_ERR_HIP, "sample readback failed");
    }
    c->device_bytes -= (long)(sizeof(double) * c->n_samples);
    dfreet(tmp);
    return rc;
}

namespace {

// The 7^3 Blackman-Harris weights of a "single" q-group, unnormalised, in
// cell units: r_max = 2*3+4 cells, r0 = r_max/2, negative values set to 0
// (q_group_mod.f90:216-239; the norm it sums is never applied).  w[(k+3)*49
// + (j+3)*7 + (i+3)].
void single_kernel(double w[343]) {
    const double a[4] = {0.353222222, -0.488, 0.145, -0.010222222};
    const double r_max = 10.0, r0 = r_max / 2.0;
    for (int k = -3; k <= 3; ++k)
        for (int j = -3; j <= 3; ++j)
            for (int i = -3; i <= 3; ++i) {
                const double x = i, y = j, z = k;
                const double r = std::sqrt(x * x + y * y + z * z);
                double v = a[0] + a[1] * std::cos(2.0 * kPi * (r - r0) / r_max) +
                           a[2] * std::cos(2.0 * kPi * 2.0 * (r - r0) / r_max) +
                           a[3] * std::cos(2.0 * kPi * 3.0 * (r - r0) / r_max);
                if (v < 0.0) v = 0.0;
                w[(k + 3) * 49 + (j + 3) * 7 + (i + 3)] = v;
            }
}

}  // namespace

int omx3d_set_q_groups(omx3d_ctx* c, int32_t n_groups, const int32_t* group_type, int32_t n, const int32_t* group_of,
                       const int32_t* ijk, double density, double dt_q) {
    CHECK_CTX(c);
    if (c->nq) return fail(OMX_ERR_STATE, "quantum sites already set");
    if (n <= 0 || !ijk) return fail(OMX_ERR_ARG, "bad site list");
    if (n_groups <= 0 || !group_type || (n_groups > 1 && !group_of)) return fail(OMX_ERR_ARG, "bad group list");
    if (!(dt_q > 0.0)) return fail(OMX_ERR_ARG, "dt_q must be positive");
    for (int g = 0; g < n_groups; ++g)
        if (group_type[g] != OMX_Q_MATERIAL && group_type[g] != OMX_Q_SINGLE)
            return fail(OMX_ERR_ARG, "q-group type must be OMX_Q_MATERIAL or OMX_Q_SINGLE");
    const int nn[3] = {c->d.nx, c->d.ny, c->d.nz};
    double wk[343];
    single_kernel(wk);
    // the cells each system writes, in the reference's order (systems in
    // group order; a kernel k, then j, then i fastest)
    struct Entry { int cell[3]; int sys, grp; double coef; };
    std::vector<Entry> ent;
    std::vector<long> q_off(n);
    int last_g = -1;
    for (int q = 0; q < n; ++q) {
        const int g = n_groups > 1 ? group_of[q] : 0;
        if (g < 0 || g >= n_groups || g < last_g) return fail(OMX_ERR_ARG, "systems must be listed in group order");
        last_g = g;
        const int v[3] = {ijk[q], ijk[q + n], ijk[q + 2 * n]};
        for (int a = 0; a < 3; ++a)
            if (v[a] < 1 || v[a] > nn[a]) return fail(OMX_ERR_RANGE, "site index outside the grid");
        q_off[q] = c->L.at(v[0], v[1], v[2]);
        if (group_type[g] == OMX_Q_MATERIAL) {
            ent.push_back({{v[0], v[1], v[2]}, q, g, density});
            continue;
        }
        for (int k = -3; k <= 3; ++k)
            for (int j = -3; j <= 3; ++j)
                for (int i = -3; i <= 3; ++i) {
                    const int u[3] = {v[0] + i, v[1] + j, v[2] + k};
                    for (int a = 0; a < 3; ++a)
                        if (u[a] < 1 || u[a] > nn[a])
                            return fail(OMX_ERR_RANGE, "a single q-group's 7^3 kernel reaches outside the grid");
                    ent.push_back({{u[0], u[1], u[2]}, q, g, density * wk[(k + 3) * 49 + (j + 3) * 7 + (i + 3)]});
                }
    }
    // the last group's scatter copies J_old <- J over the whole grid, so only
    // its own writes leave a dJ (an empty last group leaves none)
    const int G = n_groups - 1;
    std::map<std::tuple<int, int, int>, int> slot_of;
    std::vector<std::tuple<int, int, int>> slot_cell;
    std::vector<QSlot> slot;
    for (const Entry& e : ent) {
        auto key = std::make_tuple(e.cell[0], e.cell[1], e.cell[2]);
        auto it = slot_of.find(key);
        int s;
        if (it == slot_of.end()) {
            s = (int)slot.size();
            slot_of[key] = s;
            slot_cell.push_back(key);
            slot.push_back({-1, -1, 1, 0, 0.0, 0.0});
        } else {
            s = it->second;
        }
        QSlot& S = slot[s];
        if (e.grp < G) {        // an earlier group: the value the last group's J_old will see
            S.prev = e.sys; S.coef_prev = e.coef;
        }
        S.sys = e.sys; S.coef = e.coef;
        S.mode = e.grp == G ? 0 : 1;
    }
    const int ns = (int)slot.size();
    int rc = dalloc(c, (void**)&c->d_q_off, sizeof(long) * n);
    if (!rc) rc = dalloc(c, (void**)&c->d_q_slot, sizeof(QSlot) * ns);
    if (!rc) rc = dalloc(c, (void**)&c->d_J, sizeof(double) * 9 * ns);
    if (!rc) rc = dalloc(c, (void**)&c->d_q_e, sizeof(double) * 3 * n);
    if (!rc) rc = dalloc(c, (void**)&c->d_q_dpdt, sizeof(double) * 3 * n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->d_q_off, q_off.data(), sizeof(long) * n, hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemcpyAsync(c->d_q_slot, slot.data(), sizeof(QSlot) * ns, hipMemcpyHostToDevice, c->s));
    // cells whose E component sees J: the site cell (term i) and the cell one
    // below along the component axis (term i+1)
    for (int comp = 0; comp < 3; ++comp) {
        std::map<std::tuple<int, int, int>, std::pair<int, int>> aff;
        for (int s = 0; s < ns; ++s) {
            int v[3] = {std::get<0>(slot_cell[s]), std::get<1>(slot_cell[s]), std::get<2>(slot_cell[s])};
            aff[std::make_tuple(v[0], v[1], v[2])];
            v[comp] -= 1;
            if (v[comp] >= 1) aff[std::make_tuple(v[0], v[1], v[2])];
        }
        std::vector<long> o;
        std::vector<int> A, B;
        for (auto& kv : aff) {
            int v[3] = {std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)};
            auto ia = slot_of.find(kv.first);
            int w[3] = {v[0], v[1], v[2]};
            w[comp] += 1;
            auto ib = slot_of.find(std::make_tuple(w[0], w[1], w[2]));
            o.push_back(c->L.at(v[0], v[1], v[2]));
            A.push_back(ia == slot_of.end() ? -1 : ia->second);
            B.push_back(ib == slot_of.end() ? -1 : ib->second);
        }
        c->n_aff[comp] = (int)o.size();
        c->h_aff_ijk[comp].clear();
        for (auto& kv : aff) {
            c->h_aff_ijk[comp].push_back(std::get<0>(kv.first));
            c->h_aff_ijk[comp].push_back(std::get<1>(kv.first));
            c->h_aff_ijk[comp].push_back(std::get<2>(kv.first));
        }
        rc = dalloc(c, (void**)&c->d_aff_off[comp], sizeof(long) * o.size());
        if (!rc) rc = dalloc(c, (void**)&c->d_aff_a[comp], sizeof(int) * A.size());
        if (!rc) rc = dalloc(c, (void**)&c->d_aff_b[comp], sizeof(int) * B.size());
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(c->d_aff_off[comp], o.data(), sizeof(long) * o.size(), hipMemcpyHostToDevice, c->s));
        HIPCHK(hipMemcpyAsync(c->d_aff_a[comp], A.data(), sizeof(int) * A.size(), hipMemcpyHostToDevice, c->s));
        HIPCHK(hipMemcpyAsync(c->d_aff_b[comp], B.data(), sizeof(int) * B.size(), hipMemcpyHostToDevice, c->s));
    }
    HIPCHK(hipStreamSynchronize(c->s));
    c->nq = n;
    c->nslot = ns;
    c->q_density = density;
    c->q_groups = n_groups;
    // mxim_mxll.f90:73-79: n_skip_steps = NINT(dt_q/dt) when dt_q > dt, else
    // 1; the quantum step is then n_skip_steps * dt
    c->n_skip = dt_q > c->d.dt ? (int)std::lround(dt_q / c->d.dt) : 1;
    c->q_dt = (double)c->n_skip * c->d.dt;
    c->tq_step_old = 0;
    c->patches_built = false;
    return OMX_OK;
}

int omx3d_set_q_sites(omx3d_ctx* c, int32_t n, const int32_t* ijk, double density, double dt_q) {
    const int32_t material = OMX_Q_MATERIAL;
    return omx3d_set_q_groups(c, 1, &material, n, nullptr, ijk, density, dt_q);
}

int omx3d_gather_e(omx3d_ctx* c, double* e) {
    CHECK_CTX(c);
    if (!c->nq) return fail(OMX_ERR_STATE, "no quantum sites");
    if (c->prec == 4)
        hipLaunchKernelGGL(k_gather_e<float>, dim3(blocks_for(c->nq, 256)), dim3(256), 0, c->s, make_args<float>(c),
                           c->d_q_off, c->nq, c->d_q_e);
    else
        hipLaunchKernelGGL(k_gather_e<double>, dim3(blocks_for(c->nq, 256)), dim3(256), 0, c->s, make_args<double>(c),
                           c->d_q_off, c->nq, c->d_q_e);
    HIPCHK(hipGetLastError());
    std::vector<double> h(3 * (size_t)c->nq);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_q_e, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (e)
        for (int q = 0; q < c->nq; ++q)
            for (int a = 0; a < 3; ++a) e[q + (long)c->nq * a] = h[3 * q + a];
    return OMX_OK;
}

int omx3d_scatter_j(omx3d_ctx* c, const double* dpdt) {
    CHECK_CTX(c);
    if (!c->nq) return fail(OMX_ERR_STATE, "no quantum sites");
    if (!dpdt) return fail(OMX_ERR_ARG, "null dPt/dt");
    std::vector<double> h(3 * (size_t)c->nq);
    for (int q = 0; q < c->nq; ++q)
        for (int a = 0; a < 3; ++a) h[3 * q + a] = dpdt[q + (long)c->nq * a];
    HIPCHK(hipMemcpyAsync(c->d_q_dpdt, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->s));
    hipLaunchKernelGGL(k_scatter_j, dim3(blocks_for(c->nslot, 256)), dim3(256), 0, c->s, c->d_q_slot, c->nslot,
                       c->d_q_dpdt, c->q_dt, c->d_J);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    return OMX_OK;
}

int omx3d_set_two_level(omx3d_ctx* c, const double* omega0, const double* mu, const double* pol,
                        const double* gamma2) {
    CHECK_CTX(c);
    if (!c->nq) return fail(OMX_ERR_STATE, "set quantum sites first");
    if (!omega0 || !mu || !pol || !gamma2) return fail(OMX_ERR_ARG, "null emitter parameter");
    const int n = c->nq;
    std::vector<double> st(16 * (size_t)n, 0.0);
    for (int q = 0; q < n; ++q) {
        double* s = &st[16 * (size_t)q];
        s[0] = omega0[q]; s[1] = mu[q];
        const double px = pol[q], py = pol[q + n], pz = pol[q + 2 * n];
        const double nr = std::sqrt(px * px + py * py + pz * pz);
        if (!(nr > 0.0)) return fail(OMX_ERR_ARG, "zero emitter polarisation");
        s[2] = px / nr; s[3] = py / nr; s[4] = pz / nr;
        s[5] = gamma2[q];
        s[6] = 1.0;  // ground state
    }
    if (!c->d_tls) {
        int rc = dalloc(c, (void**)&c->d_tls, sizeof(double) * st.size());
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(c->d_tls, st.data(), sizeof(double) * st.size(), hipMemcpyHostToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->tls = true;
    return OMX_OK;
}

int omx3d_two_level_step(omx3d_ctx* c) {
    CHECK_CTX(c);
    if (!c->tls) return fail(OMX_ERR_STATE, "no two-level emitters");
    return c->prec == 4 ? launch_two_level<float>(c) : launch_two_level<double>(c);
}

int omx3d_get_two_level(omx3d_ctx* c, double* dipole, double* rho) {
    CHECK_CTX(c);
    if (!c->tls) return fail(OMX_ERR_STATE, "no two-level emitters");
    const int n = c->nq;
    std::vector<double> st(16 * (size_t)n);
    HIPCHK(hipMemcpyAsync(st.data(), c->d_tls, sizeof(double) * st.size(), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    for (int q = 0; q < n; ++q) {
        const double* s = &st[16 * (size_t)q];
        for (int a = 0; a < 3; ++a) if (dipole) dipole[q + (long)n * a] = s[10 + a];
        for (int a = 0; a < 4; ++a) if (rho) rho[q + (long)n * a] = s[6 + a];
    }
    return OMX_OK;
}

int omx3d_copy_field(omx3d_ctx* c, int32_t field, double* host, int32_t to_device) {
    CHECK_CTX(c);
    if (field < 1 || field > 6 || !host) return fail(OMX_ERR_ARG, "bad field or null buffer");
    return c->prec == 4 ? copy_field_impl<float>(c, field, host, to_device)
                        : copy_field_impl<double>(c, field, host, to_device);
}

int omx3d_energy(omx3d_ctx* c, double* w64, double* w32c) {
    CHECK_CTX(c);
    return c->prec == 4 ? energy_impl<float>(c, w64, w32c) : energy_impl<double>(c, w64, w32c);
}

int omx3d_energy_noalloc(omx3d_ctx* c, double* w_dev) {
    CHECK_CTX(c);
    if (!w_dev) return fail(OMX_ERR_ARG, "null output");
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof(attr));
    if (hipPointerGetAttributes(&attr, w_dev) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        return fail(OMX_ERR_ARG, "the output of omx3d_energy_noalloc must be device memory (2 doubles)");
    }
    if (attr.device != c->device)
        return fail(OMX_ERR_ARG, "the output must live on the context's GPU");
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != c->device) HIPCHK(hipSetDevice(c->device));
    double* part = nullptr;
    int rc = scratch_on(c->device, &part);
    if (!rc) rc = c->prec == 4 ? launch_energy<float>(c, part, w_dev) : launch_energy<double>(c, part, w_dev);
    if (cur != c->device) {
        const hipError_t e = hipSetDevice(cur);
        if (!rc && e != hipSuccess) rc = fail(OMX_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    }
    return rc;
}

int omx3d_get_tfsf_line(omx3d_ctx* c, double* ex1, double* hy1, int32_t* kb, int32_t* n) {
    CHECK_CTX(c);
    if (!c->tfsf) return fail(OMX_ERR_STATE, "no TF/SF plane wave");
    if (kb) *kb = c->tf_kb;
    if (n) *n = c->tf_n;
    if (ex1) HIPCHK(hipMemcpyAsync(ex1, c->d_tf, sizeof(double) * c->tf_n, hipMemcpyDeviceToHost, c->s));
    if (hy1) HIPCHK(hipMemcpyAsync(hy1, c->d_tf + c->tf_n, sizeof(double) * c->tf_n, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    return OMX_OK;
}

int omx3d_set_tfsf(omx3d_ctx* c, const int32_t lo[3], const int32_t hi[3], double e0, double w0, double t0,
                   double tau, int32_t max_steps) {
    CHECK_CTX(c);
    if (!lo || !hi) return fail(OMX_ERR_ARG, "null box");
    const int n[3] = {c->d.nx, c->d.ny, c->d.nz};
    for (int a = 0; a < 3; ++a) {
        const int guard = c->d.boundaries[a] == OMX_BC_CPML ? c->d.npml + 1 : 2;
        if (lo[a] < guard || hi[a] > n[a] - guard || hi[a] - lo[a] < 2)
            return fail(OMX_ERR_RANGE, "TF/SF box must lie inside the non-CPML interior with at least 2 cells per axis");
    }
    if (max_steps < 1) return fail(OMX_ERR_ARG, "max_steps must be >= 1");
    for (int a = 0; a < 3; ++a) { c->tf_lo[a] = lo[a]; c->tf_hi[a] = hi[a]; }
    // 1D line long enough that nothing reflected at its ends reaches the box
    // within max_steps (the pulse moves c*dt/dr = 1/2 cell per step).
    const int pad = max_steps / 2 + 16;
    c->tf_ks = lo[2] - 3;
    c->tf_kb = c->tf_ks - pad;
    const int ke = hi[2] + pad;
    c->tf_n = ke - c->tf_kb + 1;
    c->tf_e0 = e0; c->tf_w0 = w0; c->tf_t0 = t0; c->tf_tau = tau;
    dfreet(c->d_tf);
    int rc = dalloc(c, (void**)&c->d_tf, sizeof(double) * 2 * c->tf_n);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->s));
    c->tfsf = true;
    c->tf_fold_e = -1;
    c->patches_built = false;
    return OMX_OK;
}

int omx3d_run(omx3d_ctx* c, int32_t tt_first, int32_t n_steps, int32_t record_every) {
    CHECK_CTX(c);
    if (n_steps < 0) return fail(OMX_ERR_ARG, "n_steps must be >= 0");
    return c->prec == 4 ? run_impl<float>(c, tt_first, n_steps, record_every)
                        : run_impl<double>(c, tt_first, n_steps, record_every);
}

int omx3d_set_step_mode(omx3d_ctx* c, int32_t mode, int32_t rows, int32_t xchunk) {
    CHECK_CTX(c);
    if (mode != 0 && mode != 1) return fail(OMX_ERR_ARG, "step mode must be 0 (fused where it applies) or 1");
    if (rows != 0 && rows != 1 && rows != 2) return fail(OMX_ERR_ARG, "fused rows must be 1 or 2");
    if (xchunk < 0) return fail(OMX_ERR_ARG, "xchunk must be >= 0");
    c->step_mode = mode;
    if (rows) c->wg_rows = rows;
    if (xchunk) c->wg_xchunk = xchunk;
    return OMX_OK;
}

int omx3d_get_fused_shape(omx3d_ctx* c, int32_t* rows, int32_t* xchunk, int32_t* ny_waves) {
    CHECK_CTX(c);
    int nzw = 0, nyw = 0;
    wg_shape(c, nzw, nyw);
    if (rows) *rows = c->wg_rows;
    if (xchunk) *xchunk = c->wg_xchunk;
    if (ny_waves) *ny_waves = nyw;
    return OMX_OK;
}

int omx3d_set_fused_waves(omx3d_ctx* c, int32_t ny_waves) {
    CHECK_CTX(c);
    if (ny_waves < 0 || ny_waves > 16) return fail(OMX_ERR_ARG, "wavefronts along y must be 0..16");
    c->wg_ny = ny_waves;
    return OMX_OK;
}

int omx3d_get_step_mode(omx3d_ctx* c, int32_t* fused_last_run) {
    CHECK_CTX(c);
    if (fused_last_run) *fused_last_run = c->last_fused ? 1 : 0;
    return OMX_OK;
}

int omx3d_set_profiling(omx3d_ctx* c, int32_t every) {
    CHECK_CTX(c);
    if (every < 0) return fail(OMX_ERR_ARG, "profiling stride must be >= 0");
    c->prof_every = every;
    c->prof_ms_h = c->prof_ms_e = 0.0;
    c->prof_n = 0;
    return OMX_OK;
}

int omx3d_get_profile(omx3d_ctx* c, double* ms_h, double* ms_e, int32_t* n) {
    CHECK_CTX(c);
    if (ms_h) *ms_h = c->prof_ms_h;
    if (ms_e) *ms_e = c->prof_ms_e;
    if (n) *n = c->prof_n;
    return OMX_OK;
}

int omx3d_set_source_patch(omx3d_ctx* c, int32_t radius) {
    CHECK_CTX(c);
    if (radius < 0 || radius > 64) return fail(OMX_ERR_ARG, "patch radius must be in 0..64");
    HIPCHK(hipStreamSynchronize(c->s));
    if (c->patch_cells) {
        // fields hold the rounded shadows already; the new boxes start from them
        free_patches(c);
    }
    c->patch_radius = radius;
    c->patches_built = false;
    return OMX_OK;
}

int omx3d_get_source_patch(omx3d_ctx* c, int32_t* radius, int32_t* n_boxes, int64_t* n_cells) {
    CHECK_CTX(c);
    if (!c->patches_built) {
        HIPCHK(hipStreamSynchronize(c->s));
        int rc = build_patches(c);
        if (rc) return rc;
    }
    if (radius) *radius = c->patch_radius;
    if (n_boxes) *n_boxes = (int32_t)c->h_boxes.size();
    if (n_cells) {   // the boxes' cells (the storage pads each box to kPatchAlign)
        long n = 0;
        for (const PatchBox& B : c->h_boxes) n += (long)B.w[0] * B.w[1] * B.w[2];
        *n_cells = n;
    }
    return OMX_OK;
}

int omx3d_synchronize(omx3d_ctx* c) {
    CHECK_CTX(c);
    HIPCHK(hipStreamSynchronize(c->s));
    return OMX_OK;
}

int omx3d_device_bytes(omx3d_ctx* c, int64_t* bytes) {
    CHECK_CTX(c);
    if (bytes) *bytes = c->device_bytes;
    return OMX_OK;
}

int omxb_create(int32_t n, omx3d_ctx* const* grids, omxb_ctx** out) {
    if (!out || !grids || n < 1) return fail(OMX_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    omxb_ctx* b = new omxb_ctx();
    b->n = n;
    b->prec = grids[0] ? grids[0]->prec : 8;
    for (int q = 0; q < n; ++q) {
        omx3d_ctx* c = grids[q];
        if (!c) { delete b; return fail(OMX_ERR_ARG, "null grid in batch"); }
        if (c->prec != b->prec) { delete b; return fail(OMX_ERR_ARG, "batch grids must share one precision"); }
        if (c->rows_y > 4) { delete b; return fail(OMX_ERR_ARG, "batch grids need rows_y <= 4"); }
        if (c->tfsf || c->nq) { delete b; return fail(OMX_ERR_ARG, "batched grids take no TF/SF or quantum sites"); }
        for (int a = 0; a < 3; ++a) if (c->d.boundaries[a] == OMX_BC_PERIODIC) b->periodic = true;
        b->g.push_back(c);
    }
    if (hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess) { delete b; return fail(OMX_ERR_HIP, "stream"); }
    if (hipMalloc((void**)&b->d_step, sizeof(long)) != hipSuccess ||
        hipMalloc((void**)&b->d_overflow, sizeof(int)) != hipSuccess) {
        omxb_destroy(b);
        return fail(OMX_ERR_NOMEM, "batch allocation failed");
    }
    for (auto* c : b->g) (void)hipStreamSynchronize(c->s);
    int rc = b->prec == 4 ? batch_setup<float>(b) : batch_setup<double>(b);
    if (rc) { omxb_destroy(b); return rc; }
    *out = b;
    return OMX_OK;
}

int omxb_run(omxb_ctx* b, int32_t tt_first, int32_t n_steps, int32_t record_every, int32_t use_graph) {
    if (!b) return fail(OMX_ERR_ARG, "null batch");
    if (n_steps < 0) return fail(OMX_ERR_ARG, "n_steps must be >= 0");
    return b->prec == 4 ? batch_run<float>(b, tt_first, n_steps, record_every, use_graph)
                        : batch_run<double>(b, tt_first, n_steps, record_every, use_graph);
}

int omxb_destroy(omxb_ctx* b) {
    if (!b) return OMX_OK;
    if (b->s) (void)hipStreamSynchronize(b->s);
    dfree(b->d_args); dfreet(b->d_tile_start);
    dfreet(b->d_src_field); dfreet(b->d_src_off); dfreet(b->d_src_ka); dfreet(b->d_src_kb); dfreet(b->d_src_of);
    dfreet(b->d_sp); dfreet(b->d_br); dfreet(b->d_grid_of); dfreet(b->d_local); dfreet(b->d_code); dfreet(b->d_off);
    dfreet(b->d_overflow); dfreet(b->d_step); dfreet(b->d_disp); dfreet(b->d_disp_grid);
    dfreet(b->d_pbox); dfreet(b->d_src_sh);
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
    return OMX_OK;
}

// Times the tile shapes (rows_y in {1,2,4} x xchunk in {1,2,4,8,16} x tile
// order 0/1/2) with `steps` H+E steps each and keeps the fastest.  The fields,
// CPML and dispersive state are cleared afterwards, so call it before a run.
int omx3d_autotune_step(omx3d_ctx* c, int32_t steps, int32_t* mode, double* ms_separate, double* ms_fused) {
    CHECK_CTX(c);
    if (steps < 1) return fail(OMX_ERR_ARG, "steps must be >= 1");
    int rc = check_sources_dispersive(c);
    if (rc) return rc;
    rc = tfsf_fold_check(c);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // times `steps` steps of one kind after two untimed ones (fields as they
    // come: the state is cleared at the end)
    auto timed = [&](bool fused, double& ms) -> int {
        const bool fold = !fused && c->tfsf, fold_e = fold && c->tf_fold_e == 1;
        const bool fold_wg = fused && c->tfsf && c->tf_fold_e == 1;
        for (int t = 0; t < steps + 2; ++t) {
            if (t == 2) HIPCHK(hipEventRecord(e0, c->s));
            int r;
            if (fused) {
                r = c->prec == 4 ? launch_fused<float>(c, nullptr, nullptr, fold_wg, 0.0)
                                 : launch_fused<double>(c, nullptr, nullptr, fold_wg, 0.0);
            } else {
                r = c->prec == 4 ? launch_h<float>(c, fold) : launch_h<double>(c, fold);
                if (!r)
                    r = c->prec == 4 ? launch_e<float>(c, 0.0, 0.0, fold_e, 0.0)
                                     : launch_e<double>(c, 0.0, 0.0, fold_e, 0.0);
            }
            if (!r && !fold_e && !fold_wg)
                r = c->prec == 4 ? launch_tfsf_e<float>(c, 0.0) : launch_tfsf_e<double>(c, 0.0);
            if (r) return r;
        }
        HIPCHK(hipEventRecord(e1, c->s));
        HIPCHK(hipEventSynchronize(e1));
        float f = 0.f;
        HIPCHK(hipEventElapsedTime(&f, e0, e1));
        ms = f / steps;
        return OMX_OK;
    };
    const int mode0 = c->step_mode, r0 = c->wg_rows, x0 = c->wg_xchunk, y0 = c->wg_ny;
    double t_sep = 0.0, t_fus = -1.0;
    rc = timed(false, t_sep);
    int best_r = r0, best_x = x0, best_y = y0;
    c->step_mode = 0;
    if (!rc && fused_applies(c)) {
        // fused tile shapes: rows per wavefront, planes per chunk, wavefronts along y
        // (small workgroups share a CU with others, large ones recompute fewer
        // halo rows: the best trade differs by grid content)
        const int cand32[6][3] = {{2, 32, 0}, {2, 16, 0}, {1, 32, 0}, {2, 32, 2}, {2, 16, 1}, {2, 8, 1}};
        const int cand64[5][3] = {{1, 32, 0}, {1, 16, 0}, {2, 32, 0}, {1, 32, 2}, {1, 16, 1}};
        const int nc = c->prec == 4 ? 6 : 5;
        for (int q = 0; q < nc && !rc; ++q) {
            const int* cd = c->prec == 4 ? cand32[q] : cand64[q];
            c->wg_rows = cd[0]; c->wg_xchunk = cd[1]; c->wg_ny = cd[2];
            double t = 0.0;
            rc = timed(true, t);
            if (!rc && (t_fus < 0.0 || t < t_fus)) { t_fus = t; best_r = cd[0]; best_x = cd[1]; best_y = cd[2]; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) { c->step_mode = mode0; c->wg_rows = r0; c->wg_xchunk = x0; c->wg_ny = y0; return rc; }
    // the fused step replaces the default only for a clear gain (a margin over
    // the timing noise of a few steps)
    const bool use_fused = t_fus >= 0.0 && t_fus < 0.98 * t_sep;
    c->step_mode = use_fused ? 0 : 1;
    c->wg_rows = best_r; c->wg_xchunk = best_x; c->wg_ny = best_y;
    if (!use_fused) {
        // the separate kernels do not need the fused step's second buffers
        HIPCHK(hipStreamSynchronize(c->s));
        for (auto& p : c->f_alt) dfree(p);
        for (auto& p : c->psi_h_alt) dfree(p);
    }
    rc = reset_run_state(c);
    if (rc) return rc;
    if (mode) *mode = c->step_mode;
    if (ms_separate) *ms_separate = t_sep;
    if (ms_fused) *ms_fused = t_fus;
    return OMX_OK;
}

int omx3d_autotune(omx3d_ctx* c, int32_t steps, int32_t* rows_y, int32_t* xchunk, int32_t* tile_order,
                   double* ms_per_step) {
    CHECK_CTX(c);
    if (steps < 1) return fail(OMX_ERR_ARG, "steps must be >= 1");
    // *rows_y of 1, 2 or 4 on entry: keep that row count (a batch runs all
    // its grids with one workgroup height), tune planes and order only
    const int fixed = rows_y ? *rows_y : 0;
    if (fixed != 0 && fixed != 1 && fixed != 2 && fixed != 4) return fail(OMX_ERR_ARG, "rows_y must be 0, 1, 2 or 4");
    int rows_c[3] = {1, 2, 4};
    if (fixed) rows_c[0] = rows_c[1] = rows_c[2] = fixed;
    const int xc_c[5] = {1, 2, 4, 8, 16};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e30;
    int br = c->rows_y, bx = c->xchunk, bo = c->tile_order;
    for (int order = 0; order < 3; ++order)
    for (int ri = 0; ri < (fixed ? 1 : 3); ++ri)
        for (int x : xc_c) {
            const int r = rows_c[ri];
            c->rows_y = r; c->xchunk = x; c->tile_order = order;
            for (int w = 0; w < 2; ++w) {   // warm-up launch pair
                int rc = c->prec == 4 ? launch_h<float>(c, false) : launch_h<double>(c, false);
                if (!rc)
                    rc = c->prec == 4 ? launch_e<float>(c, 0.0, 0.0, false, 0.0) : launch_e<double>(c, 0.0, 0.0, false, 0.0);
                if (rc) return rc;
            }
            HIPCHK(hipEventRecord(e0, c->s));
            for (int t = 0; t < steps; ++t) {
                int rc = c->prec == 4 ? launch_h<float>(c, false) : launch_h<double>(c, false);
                if (!rc)
                    rc = c->prec == 4 ? launch_e<float>(c, 0.0, 0.0, false, 0.0) : launch_e<double>(c, 0.0, 0.0, false, 0.0);
                if (rc) return rc;
            }
            HIPCHK(hipEventRecord(e1, c->s));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms / steps < best) { best = ms / steps; br = r; bx = x; bo = order; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->rows_y = br; c->xchunk = bx; c->tile_order = bo;
    int rc = reset_run_state(c);
    if (rc) return rc;
    if (rows_y) *rows_y = br;
    if (xchunk) *xchunk = bx;
    if (tile_order) *tile_order = bo;
    if (ms_per_step) *ms_per_step = best;
    return OMX_OK;
}

}  // extern "C"

// The 2D and 1D solvers.
#include "omx_lowdim.hpp"
