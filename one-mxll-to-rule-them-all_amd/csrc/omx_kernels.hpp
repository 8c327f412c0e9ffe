// omx_kernels.hpp -- gfx950 kernels of libomx_fdtd_hip.
//
// Every stencil kernel works on the private layout of omx::Layout (z fastest).
// A workgroup is 64 lanes along z times `rows` rows along y and marches a
// chunk of x planes, carrying the x-neighbour values in registers; the y and
// z neighbours are re-read from L1/L2.  Workgroup ids are remapped so that
// consecutive tiles (which share halo rows) land on the same XCD.
#pragma once

#include <type_traits>

#include "omx_internal.hpp"

namespace omx {

constexpr int kLanes = 64;

// Bijective map from the hardware workgroup id (dispatched round-robin over
// the 8 XCDs of an MI355X) to a tile id such that each XCD receives one
// contiguous run of tiles.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int per = nb >> 3, rem = nb & 7;
    const int x = b & 7, q = b >> 3;
    return x * per + (x < rem ? x : rem) + q;
}

// Tile of workgroup b of nb tile workgroups (tiles numbered z tile fastest,
// then row block, then x chunk; ntz z tiles per row block).  The dispatcher
// deals consecutive workgroups round-robin to the 8 XCDs.  Order 0 keeps
// the block order: consecutive workgroups cover consecutive memory, so the
// whole chip streams one region of every array at a time, and the halo
// lines of the neighbouring z tile and row, read by another XCD, come from
// the memory-side cache.  Order 1 gives every XCD one contiguous slab of
// tiles, which keeps the halos in its own L2 but streams eight regions at
// once.  Order 2 streams like order 0 but deals whole row blocks: the z
// tiles of a row block run on one XCD, whose L2 then serves their shared
// k-1 / k+1 halo lines.
__device__ __forceinline__ int tile_of_block(int order, int b, int nb, int ntz) {
    if (order == 1) return xcd_tile(b, nb);
    if (order == 2) {
        const int G = 8 * ntz;                 // a group: 8 row blocks of ntz z tiles
        if (b >= nb / G * G) return b;         // the tail keeps the block order
        const int g = b / G, r = b - g * G;
        return ((g << 3) + (r & 7)) * ntz + (r >> 3);
    }
    return b;
}

// Field access by 32-bit byte offset from a uniform base pointer (callers
// take these paths only when every array of the grid is below 4 GiB).
// The device arrays all live in global memory; stating it lets the
// compiler use global (not flat) loads and stores also where a pointer comes
// from memory (the argument blocks of a batch), with a scalar base and a
// 32-bit vector offset.
#define OMX_GLOBAL __attribute__((address_space(1)))
// (C-style casts: the address-space conversion of a generic pointer)
template <typename T>
__device__ __forceinline__ T ldb(const T* p, uint32_t b) {
    return *(const OMX_GLOBAL T*)((const OMX_GLOBAL char*)p + b);
}
template <typename T>
__device__ __forceinline__ void stb(T* p, uint32_t b, T v) {
    *(OMX_GLOBAL T*)((OMX_GLOBAL char*)p + b) = v;
}
// Element i of a global array.
template <typename T>
__device__ __forceinline__ T gld(const T* p, long i) {
    return ((const OMX_GLOBAL T*)p)[i];
}
template <typename T>
__device__ __forceinline__ void gst(T* p, long i, T v) {
    ((OMX_GLOBAL T*)p)[i] = v;
}
// Uniform read of a table the kernel never writes (scalar load).  The tiles
// read their per-plane x coefficients this way: a scalar load does not wait
// on the vector memory counter, so a load the compiler places after a
// branch costs a scalar-cache round trip, not a memory round trip.
template <typename T>
__device__ __forceinline__ T ldc(const T* p, int i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Offsets of a cell into the field, psi and tag arrays of the H/E tiles.
// B32: 32-bit byte offsets (every array of the grid below 4 GiB), so that a
// load takes a scalar base and one 32-bit vector offset with no 64-bit
// address arithmetic, which leaves vector registers for loads in flight;
// otherwise 64-bit element indices.
template <typename T, bool B32>
struct Off {
    using I = long;
    static __device__ __forceinline__ I of(long e) { return e; }
    static __device__ __forceinline__ T ld(const T* p, I o) { return gld(p, o); }
    static __device__ __forceinline__ void st(T* p, I o, T v) { gst(p, o, v); }
    static __device__ __forceinline__ int tag(const uint8_t* p, I o) { return gld(p, o); }
};
template <typename T>
struct Off<T, true> {
    using I = uint32_t;
    static __device__ __forceinline__ I of(long e) { return (uint32_t)(e * (long)sizeof(T)); }
    static __device__ __forceinline__ T ld(const T* p, I o) { return ldb(p, o); }
    static __device__ __forceinline__ void st(T* p, I o, T v) { stb(p, o, v); }
    static __device__ __forceinline__ int tag(const uint8_t* p, I o) { return gld(p, (long)(o / (uint32_t)sizeof(T))); }
};

// ---------------------------------------------------------------------------
// K8: FP64 coefficient precompute on the device.
// CPML grading follows init_3Dgrid (src/FDTD/mxll_3D_mod.f90:294-318): the
// grading ratios and their cubes are evaluated in single precision, exactly
// as the reference's default-REAL expressions are, then promoted.
// ---------------------------------------------------------------------------
__device__ inline double cube_f(float r) { return (double)((r * r) * r); }

struct CpmlProfile {
    double sig_max, alpha_scale, dt, dr;
    int npml;
};

__device__ inline void cpml_e(const CpmlProfile& p, int kk, double& sig, double& alpha, double& kappa) {
#pragma clang fp contract(off)
    const float den = (float)p.npml - 1.0f;
    const float r = (float)(p.npml - kk) / den;
    const double r3 = cube_f(r);
    sig = p.sig_max * r3;
    alpha = p.alpha_scale * (double)((float)(kk - 1) / den);
    kappa = 1.0 + (kKappaCPML - 1.0) * r3;
}

__device__ inline void cpml_h(const CpmlProfile& p, int kk, double& sig, double& alpha, double& kappa) {
#pragma clang fp contract(off)
    const float den = (float)p.npml - 1.0f;
    const float r = ((float)(p.npml - kk) - 0.5f) / den;
    const double r3 = cube_f(r);
    sig = p.sig_max * r3;
    alpha = p.alpha_scale * (double)(((float)kk - 0.5f) / den);
    kappa = 1.0 + (kKappaCPML - 1.0) * r3;
}

// One thread per 1-based index k in 1..max(nx,ny,nz,npml).
// out layout: be[0..npml] ce[..] bh[0..npml-1] ch[..] den_ex[0..nx+1] den_ey den_ez den_hx den_hy den_hz
__global__ void k_cpml_coeffs(CpmlProfile p, int nx, int ny, int nz, unsigned mask,
                              double* be, double* ce, double* bh, double* ch,
                              double* dex, double* dey, double* dez,
                              double* dhx, double* dhy, double* dhz) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int npml = p.npml;
    const bool any = mask != 0u;
    if (any && k <= npml) {
        double s, a, kap;
        cpml_e(p, k, s, a, kap);
        const double b = exp(-((s / kap + a) * p.dt) / kEps0);
        be[k] = b;
        if (s == 0.0 && a == 0.0 && k == npml) ce[k] = 0.0;
        else ce[k] = ((s * (b - 1.0)) / (s + kap * a)) / kap;
    }
    if (any && k <= npml - 1) {
        double s, a, kap;
        cpml_h(p, k, s, a, kap);
        const double b = exp(-((s / kap + a) * p.dt) / kEps0);
        bh[k] = b;
        ch[k] = ((s * (b - 1.0)) / (s + kap * a)) / kap;
    }
    const double inv = 1.0 / p.dr;
    const int nn[3] = {nx, ny, nz};
    double* de[3] = {dex, dey, dez};
    double* dh[3] = {dhx, dhy, dhz};
    for (int ax = 0; ax < 3; ++ax) {
        const int n = nn[ax];
        if (k > n) continue;
        double ve = inv, vh = inv;
        const bool lo = (mask >> (2 * ax)) & 1u, hi = (mask >> (2 * ax + 1)) & 1u;
        double s, a, kap;
        if (lo && k <= npml) { cpml_e(p, k, s, a, kap); ve = 1.0 / (kap * p.dr); }
        if (lo && k <= npml - 1) { cpml_h(p, k, s, a, kap); vh = 1.0 / (kap * p.dr); }
        if (hi && k >= n + 1 - npml && k <= n - 1) {
            cpml_e(p, n + 1 - k, s, a, kap); ve = 1.0 / (kap * p.dr);
            cpml_h(p, n - k, s, a, kap);     vh = 1.0 / (kap * p.dr);
        }
        de[ax][k] = ve;
        dh[ax][k] = vh;
    }
}

// FP32 b-1 with full relative precision: recompute from the exponent.
__global__ void k_cpml_bm1(CpmlProfile p, unsigned mask, float* be_t, float* bh_t) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (mask == 0u) return;
    double s, a, kap;
    if (k <= p.npml) { cpml_e(p, k, s, a, kap); be_t[k] = (float)expm1(-((s / kap + a) * p.dt) / kEps0); }
    if (k <= p.npml - 1) { cpml_h(p, k, s, a, kap); bh_t[k] = (float)expm1(-((s / kap + a) * p.dt) / kEps0); }
}

// ---------------------------------------------------------------------------
// K8 (media): update coefficients of the classical media, device FP64,
// following init_classical_medium (classical_medium_mod.f90:60-101).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_medium_coeffs(const omx_medium* __restrict__ in, int n, double dt, double dt_eps0,
                                omx_medium_coeffs* __restrict__ co, MedDev<T>* __restrict__ md) {
#pragma clang fp contract(off)
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > n) return;
    omx_medium_coeffs c;
    MedDev<T> d;
    for (int q = 0; q < OMX_MAX_POLES; ++q) {
        c.alpha_k[q] = c.beta_k[q] = c.gamma_k[q] = c.B1_k[q] = c.B2_k[q] = 0.0;
        d.al[q] = d.bt[q] = d.ga[q] = d.B1[q] = d.B2[q] = (T)0;
    }
    c.A1 = c.A2 = c.C1 = c.C2 = c.C3 = c.C4 = 0.0;
    c.inv_eps = 1.0;
    d.kind = OMX_MEDIUM_NONE; d.np = 0;
    d.cdt = (T)dt_eps0;
    d.C1 = d.C2 = d.C3 = d.C4 = d.A1 = d.A2 = (T)0;
    if (m > 0) {
        const omx_medium& M = in[m];
        d.kind = M.kind;
        if (M.kind == OMX_MEDIUM_DIELECTRIC) {
            c.inv_eps = 1.0 / M.eps_r;
            d.cdt = (T)(dt_eps0 * c.inv_eps);
        } else if (M.kind == OMX_MEDIUM_DRUDE) {
            const double gd = M.gamma_d * dt;
            c.A1 = (2.0 - gd) / (2.0 + gd);
            c.A2 = ((kEps0 * (M.omega_d * M.omega_d)) * dt) / (2.0 + gd);
            const double X = (M.eps_r * kEps0) / dt;
            c.C1 = (X - 0.5 * c.A2) / (X + 0.5 * c.A2);
            c.C3 = 1.0 / (X + 0.5 * c.A2);
            c.C4 = (0.5 * (c.A1 + 1.0)) / (X + 0.5 * c.A2);
            const double c1m1 = -c.A2 / (X + 0.5 * c.A2);
            const double a1m1 = -2.0 * gd / (2.0 + gd);
            d.C1 = (T)(sizeof(T) == 4 ? c1m1 : c.C1);
            d.A1 = (T)(sizeof(T) == 4 ? a1m1 : c.A1);
            d.A2 = (T)c.A2; d.C3 = (T)c.C3; d.C4 = (T)c.C4;
        } else if (M.kind == OMX_MEDIUM_DL) {
            const int np = M.n_poles;
            d.np = np;
            double S = 0.0;
            for (int q = 0; q < np; ++q) {
                const double w = M.omega_k[q], g = M.gamma_k[q];
                c.alpha_k[q] = (2.0 - (w * w) * (dt * dt)) / (1.0 + (0.5 * g) * dt);
                c.beta_k[q] = (g * dt - 2.0) / (g * dt + 2.0);
                c.gamma_k[q] = ((kEps0 * M.a_k[q]) * dt) / (g * dt + 2.0);
                S = S + c.gamma_k[q];
            }
            const double gd = M.gamma_d * dt;
            c.A1 = (2.0 - gd) / (2.0 + gd);
            c.A2 = ((kEps0 * (M.omega_d * M.omega_d)) * dt) / (2.0 + gd);
            const double X = (M.eps_r * kEps0) / dt;
            const double D = (X + 0.5 * c.A2) + 0.5 * S;
            c.C1 = (X - 0.5 * c.A2) / D;
            c.C2 = (0.5 * S) / D;
            c.C3 = 1.0 / D;
            c.C4 = (0.5 * (c.A1 + 1.0)) / D;
            for (int q = 0; q < np; ++q) {
                c.B1_k[q] = (0.5 * (1.0 + c.alpha_k[q])) / D;
                c.B2_k[q] = (0.5 * c.beta_k[q]) / D;
            }
            const bool f32 = sizeof(T) == 4;
            d.C1 = (T)(f32 ? -(c.A2 + 0.5 * S) / D : c.C1);
            d.A1 = (T)(f32 ? -2.0 * gd / (2.0 + gd) : c.A1);
            d.C2 = (T)c.C2; d.C3 = (T)c.C3; d.C4 = (T)c.C4; d.A2 = (T)c.A2;
            for (int q = 0; q < np; ++q) {
                const double w = M.omega_k[q], g = M.gamma_k[q];
                const double am2 = -((w * w) * (dt * dt) + g * dt) / (1.0 + (0.5 * g) * dt);
                const double bp1 = (2.0 * g * dt) / (g * dt + 2.0);
                d.al[q] = (T)(f32 ? am2 : c.alpha_k[q]);
                d.bt[q] = (T)(f32 ? bp1 : c.beta_k[q]);
                d.ga[q] = (T)c.gamma_k[q];
                d.B1[q] = (T)c.B1_k[q];
                d.B2[q] = (T)c.B2_k[q];
            }
        }
    }
    co[m] = c;
    md[m] = d;
}

// ---------------------------------------------------------------------------
// Media map: cell map (aux, Fortran order) -> staggered per-component map
// (read_init_media_3D, classical_medium_mod.f90:694-745), then one-byte tags.
// ---------------------------------------------------------------------------
__global__ void k_rasterize(const omx_shape* __restrict__ sh, int ns, int nx, int ny, int nz,
                            int8_t* __restrict__ aux) {
#pragma clang fp contract(off)
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)nx * ny * nz;
    if (c >= n) return;
    const int i = (int)(c % nx) + 1;
    const int j = (int)((c / nx) % ny) + 1;
    const int k = (int)(c / ((long)nx * ny)) + 1;
    const double x = i, y = j, z = k;
    int v = 0;
    for (int s = 0; s < ns; ++s) {
        const double* p = sh[s].p;
        bool in = false;
        switch (sh[s].type) {
            case OMX_SHAPE_SPHERE: {
                const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
                in = dx * dx + dy * dy + dz * dz <= p[3] * p[3];
            } break;
            case OMX_SHAPE_ELLIPSOID: {
                const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
                const double u = p[6] * dx + p[7] * dy + p[8] * dz;
                const double w = p[9] * dx + p[10] * dy + p[11] * dz;
                const double t = p[12] * dx + p[13] * dy + p[14] * dz;
                in = (u / p[3]) * (u / p[3]) + (w / p[4]) * (w / p[4]) + (t / p[5]) * (t / p[5]) <= 1.0;
            } break;
            case OMX_SHAPE_BOX:
                in = x >= p[0] && y >= p[1] && z >= p[2] && x <= p[3] && y <= p[4] && z <= p[5];
                break;
            case OMX_SHAPE_CYLINDER_Z: {
                const double dx = x - p[0], dy = y - p[1];
                in = dx * dx + dy * dy <= p[2] * p[2] && z >= p[3] && z <= p[4];
            } break;
            default: break;
        }
        if (!in) continue;
        if (sh[s].medium == 0) v = 0;
        else if (v == 0) v = sh[s].medium;
    }
    aux[c] = (int8_t)v;
}

// Per-component medium ids of cell (i,j,k), a literal restatement of the
// reference's per-medium if/else chain.  Neighbours past the grid read 0.
__device__ inline void stagger(const int8_t* __restrict__ aux, int nx, int ny, int nz, int nm,
                               int i, int j, int k, int m[3]) {
    auto A = [&](int a, int b, int c) -> int {
        if (a > nx || b > ny || c > nz) return 0;
        return aux[(long)(a - 1) + (long)nx * ((b - 1) + (long)ny * (c - 1))];
    };
    const int a0 = A(i, j, k);
    const int ax = A(i + 1, j, k), ay = A(i, j + 1, k), az = A(i, j, k + 1);
    m[0] = m[1] = m[2] = 0;
    for (int n = 1; n <= nm; ++n) {
        if (i == nx && a0 == n) m[0] = n;
        else if (j == ny && a0 == n) m[0] = n;
        else if (k == nz && a0 == n) m[0] = n;
        else {
            if ((a0 == n && ax == n) || (a0 != 0 && ax == n)) m[0] = n;
            if ((a0 == n && ay == n) || (a0 != 0 && ay == n)) m[1] = n;
            if ((a0 == n && az == n) || (a0 != 0 && az == n)) m[2] = n;
        }
    }
}

// Pass 1: presence of each (mx,my,mz) code.  Pass 2: tag = lut[code].
__global__ void k_media_codes(const int8_t* __restrict__ aux, int nx, int ny, int nz, int nm,
                              unsigned* __restrict__ present, const uint8_t* __restrict__ lut,
                              uint8_t* __restrict__ tag, Layout L, int32_t* __restrict__ map_out) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)nx * ny * nz;
    if (c >= n) return;
    const int i = (int)(c % nx) + 1;
    const int j = (int)((c / nx) % ny) + 1;
    const int k = (int)(c / ((long)nx * ny)) + 1;
    int m[3];
    stagger(aux, nx, ny, nz, nm, i, j, k, m);
    const int code = m[0] + (nm + 1) * (m[1] + (nm + 1) * m[2]);
    if (present) {
        // nearly every cell repeats a code already seen: read before the
        // atomic so that only first sightings contend for the word
        const unsigned bit = 1u << (code & 31);
        if (!(__atomic_load_n(&present[code >> 5], __ATOMIC_RELAXED) & bit)) atomicOr(&present[code >> 5], bit);
    }
    if (lut) tag[L.at(i, j, k)] = lut[code];
    if (map_out) { map_out[c] = m[0]; map_out[c + n] = m[1]; map_out[c + 2 * n] = m[2]; }
}

// ---------------------------------------------------------------------------
// H update: td_propagate_H_3Dfield (mxll_3D_mod.f90:546-694): bulk curl of E
// for i,j,k = 1..n, then the CPML psi recursions of the six faces, applied to
// each component in the reference's face order.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T psi_step(T psi, T b, T c, T d) {
    if constexpr (sizeof(T) == 4) return psi + (b * psi + c * d);  // b holds b-1
    else return b * psi + c * d;
}

// H of one cell, in place in hx/hy/hz (H^{n-1/2} in, H^{n+1/2} out): bulk
// curl of E for i,j,k in 1..n, then the CPML psi recursions of the faces the
// cell lies in, applied per component in the reference's face order (x, y,
// z).  psi is read from a.ph and written to pho where st is set (halo cells
// compute it without storing).  cx is the plane's coefficient: cdhx[i] in
// FP32, den_hx[i] in FP64.
template <typename T>
__device__ __forceinline__ void h_bulk(T dmu, T cx, T dhy, T dhz, T chy, T chz, T ezc, T ezn, T ezj, T eyc, T eyn,
                                       T eyk, T exc, T exk, T exj, T& hx, T& hy, T& hz) {
    if constexpr (sizeof(T) == 4) {
        // FP32: dt_mu0*den folded into one coefficient per axis (one rounding fewer)
        hx = hx + (ezc - ezj) * chy + (eyk - eyc) * chz;
        hy = hy + (ezn - ezc) * cx + (exc - exk) * chz;
        hz = hz + (eyc - eyn) * cx + (exj - exc) * chy;
    } else {
        hx = hx + dmu * ((ezc - ezj) * dhy + (eyk - eyc) * dhz);
        hy = hy + dmu * ((ezn - ezc) * cx + (exc - exk) * dhz);
        hz = hz + dmu * ((eyc - eyn) * cx + (exj - exc) * dhy);
    }
}

// x faces: psix_Hyx, psix_Hzx
template <typename T>
__device__ __forceinline__ void h_cpml_x(const YeeArgs<T>& a, T* const* pho, bool st, int i, long jk, T dmu,
                                         T ezc, T ezn, T eyc, T eyn, T& hy, T& hz) {
    const unsigned cp = a.cpml;
    const int npml = a.npml;
    if ((cp & 1u) && i <= npml - 1) {
        const int c = i;
        const long ps = (long)(i - 1) * a.sx + jk;
        const T p0 = psi_step(a.ph[0][ps], a.bh[c], a.chd[c], ezn - ezc);
        const T p1 = psi_step(a.ph[1][ps], a.bh[c], a.chd[c], eyc - eyn);
        if (st) { pho[0][ps] = p0; pho[1][ps] = p1; }
        hy = hy + dmu * p0; hz = hz + dmu * p1;
    }
    if ((cp & 2u) && i >= a.nx + 1 - npml && i <= a.nx - 1) {
        const int c = a.nx - i;
        const long ps = (long)(npml - 2 + c) * a.sx + jk;
        const T p0 = psi_step(a.ph[0][ps], a.bh[c], a.chd[c], ezn - ezc);
        const T p1 = psi_step(a.ph[1][ps], a.bh[c], a.chd[c], eyc - eyn);
        if (st) { pho[0][ps] = p0; pho[1][ps] = p1; }
        hy = hy + dmu * p0; hz = hz + dmu * p1;
    }
}

// y faces: psiy_Hxy, psiy_Hzy
template <typename T>
__device__ __forceinline__ void h_cpml_y(const YeeArgs<T>& a, T* const* pho, bool st, int i, int j, int k, T dmu,
                                         T ezc, T ezj, T exc, T exj, T& hx, T& hz) {
    const unsigned cp = a.cpml;
    const int npml = a.npml;
    const int nsh = 2 * (npml - 1);
    if ((cp & 4u) && j <= npml - 1) {
        const int c = j;
        const long ps = ((long)i * nsh + (j - 1)) * a.sz + k + a.off;
        const T p0 = psi_step(a.ph[2][ps], a.bh[c], a.chd[c], ezc - ezj);
        const T p1 = psi_step(a.ph[3][ps], a.bh[c], a.chd[c], exj - exc);
        if (st) { pho[2][ps] = p0; pho[3][ps] = p1; }
        hx = hx + dmu * p0; hz = hz + dmu * p1;
    }
    if ((cp & 8u) && j >= a.ny + 1 - npml && j <= a.ny - 1) {
        const int c = a.ny - j;
        const long ps = ((long)i * nsh + (npml - 2 + c)) * a.sz + k + a.off;
        const T p0 = psi_step(a.ph[2][ps], a.bh[c], a.chd[c], ezc - ezj);
        const T p1 = psi_step(a.ph[3][ps], a.bh[c], a.chd[c], exj - exc);
        if (st) { pho[2][ps] = p0; pho[3][ps] = p1; }
        hx = hx + dmu * p0; hz = hz + dmu * p1;
    }
}

// CPML face of one axis for a cell: psi slot and coefficient index (slot
// < 0: the cell is in no face of that axis).  H faces: low 1..npml-1 (slot
// n-1, coefficient n), high n_a+1-npml..n_a-1 (slot npml-2+c, c = n_a-n).  E
// faces: low 2..npml (slot n-1, coefficient n), high n_a+1-npml..n_a-1 (slot
// npml-1+c, c = n_a+1-n).  Valid when the two faces of an axis do not meet.
struct Face {
    int slot, c;
};
__device__ __forceinline__ Face face_h(bool lo_on, bool hi_on, int n, int na, int npml) {
    if (lo_on && n <= npml - 1) return {n - 1, n};
    if (hi_on && n >= na + 1 - npml && n <= na - 1) return {npml - 2 + na - n, na - n};
    return {-1, 0};
}
__device__ __forceinline__ Face face_e(bool lo_on, bool hi_on, int n, int na, int npml) {
    if (lo_on && n >= 2 && n <= npml) return {n - 1, n};
    if (hi_on && n >= na + 1 - npml && n <= na - 1) return {npml - 1 + na + 1 - n, na + 1 - n};
    return {-1, 0};
}

// z-face psi (psiz_Hxz/Hyz, psiz_Exz/Eyz).  Per side (0 low, 1 high), row j
// and plane i, the side's npml-1 cells m are contiguous, and the planes of a
// row follow one another:
//   index = ((side * sy + j) * (nx + 2) + i) * (npml - 1) + m.
// H faces: low k = 1..npml-1 (m = k-1), high k = nz+1-npml..nz-1 (m =
// nz-1-k); E faces: low k = 2..npml (m = k-2), high as for H.  The
// coefficient index is m+1 (H) and m+2 (E) on both sides.
__host__ __device__ inline long zpsi_index(int side, int j, int i, int m, int sy, int nx, int npml) {
    return (((long)side * sy + j) * (nx + 2) + i) * (npml - 1) + m;
}
// Elements of one z-face psi array.
__host__ __device__ inline long zpsi_elems(int sy, int nx, int npml) { return 2L * sy * (nx + 2) * (npml - 1); }

// Value of v in lane byte_addr/4 (ds_bpermute; both lanes active).
template <typename T>
__device__ __forceinline__ T lane_pull(int byte_addr, T v) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_ds_bpermute(byte_addr, __builtin_bit_cast(int, v)));
    } else {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_ds_bpermute(byte_addr, (int)(unsigned)b);
        const unsigned hi = (unsigned)__builtin_amdgcn_ds_bpermute(byte_addr, (int)(unsigned)(b >> 32));
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    }
}

// z-face psi of a wavefront, kept in registers a batch of planes at a time.
// The wavefront's face lanes (cells of its tile [kt0, kt1] in the low face
// [zl0, zl1] or the high face [zh0, zh1]; nf of them, ranked q = 0..nf-1 in
// lane order) are served by storage lanes: lane l < P * nf (P = nact / nf
// planes per batch, nact the tile's active lanes) holds rank l % nf of
// batch plane l / nf.  A batch is loaded and stored with one access per
// array, the face lanes read and return their values by ds_bpermute each
// plane.  (One access of nf lanes per plane and array costs the memory
// system more than its bytes; the batched access does not.)
struct ZBatch {
    int nf, P;          // face lanes of the wavefront and planes per batch (uniform)
    int q;              // this lane's face rank, -1: not a face lane
    int m;              // this face lane's cell in its side
    int sp, slane;      // storage: batch plane held by this lane (-1: none), its face lane
    long sidx;          // psi index of this storage lane in the first batch
};
__device__ __forceinline__ ZBatch z_batch(int lane, int kt0, int kt1, int zl0, int zl1, int zh0, int zh1, int klo0,
                                          int j, int i0, int nz, int sy, int nx, int npml) {
    ZBatch z;
    const int a0 = max(kt0, zl0), a1 = min(kt1, zl1), b0 = max(kt0, zh0), b1 = min(kt1, zh1);
    const int nlo = max(0, a1 - a0 + 1), nhi = max(0, b1 - b0 + 1);
    const int la = a0 - kt0, ha = b0 - kt0;
    z.nf = nlo + nhi;
    z.P = z.nf > 0 ? (kt1 - kt0 + 1) / z.nf : 0;
    z.q = -1; z.m = 0;
    if (lane >= la && lane < la + nlo) { z.q = lane - la; z.m = kt0 + lane - klo0; }
    else if (lane >= ha && lane < ha + nhi) { z.q = nlo + lane - ha; z.m = nz - 1 - (kt0 + lane); }
    z.sp = -1; z.slane = 0; z.sidx = 0;
    if (z.nf > 0 && lane < z.P * z.nf) {
        z.sp = lane / z.nf;
        const int sq = lane - z.sp * z.nf;
        z.slane = sq < nlo ? la + sq : ha + (sq - nlo);
        const int ks = kt0 + z.slane;
        const int side = sq < nlo ? 0 : 1;
        z.sidx = zpsi_index(side, j, i0 + z.sp, side == 0 ? ks - klo0 : nz - 1 - ks, sy, nx, npml);
    }
    return z;
}

// TF/SF plane wave, x-polarised, travelling along +z.  The incident Ex (at
// z = k) and Hy (at z = k + 1/2) come from a 1D Yee line with the 3D grid's
// dt and dr, so the incident wave has the 3D grid's numerical dispersion
// along z.  Total-field region: Ex i in [i0,i1-1], j in [j0,j1], k in [k0,k1];
// Hy i in [i0,i1-1], j in [j0,j1], k in [k0,k1-1]; Hz i in [i0,i1-1],
// j in [j0,j1-1], k in [k0,k1]; Ez i in [i0,i1], j in [j0,j1], k in [k0,k1-1].
struct TfsfArgs {
    int i0, i1, j0, j1, k0, k1;
    int kb, n1, ks;     // 1D line covers k = kb .. kb+n1-1; source node ks
    double dt_mu0, dt_eps0, inv_dr;
};

// TF/SF inside the separate H / E launches: the tile that updates a cell
// adds the cell's correction before storing it, and workgroups [b0, b0 + nb)
// of the launch advance the 1D line (H launch: hy1 from ex1; E launch: ex1
// from hy1 plus the source).  The tiles read only the half of the line the
// launch does not write.
struct TfsfLaunch {
    TfsfArgs t;
    double* ex1;
    double* hy1;
    double src;         // E launch: value added at node ks this step
    int on;
    int b0, nb;
};

// Whether index range [a0, a1] meets [f0, f1].
__device__ __forceinline__ bool meets(int a0, int a1, int f0, int f1) { return a0 <= f1 && a1 >= f0; }
// Whether [a0, a1] holds a cell of an H face (1..npml-1, n+1-npml..n-1) or
// an E face (2..npml, n+1-npml..n-1) of an axis.
__device__ __forceinline__ bool in_faces_h(bool lo_on, bool hi_on, int a0, int a1, int n, int npml) {
    return (lo_on && meets(a0, a1, 1, npml - 1)) || (hi_on && meets(a0, a1, n + 1 - npml, n - 1));
}
__device__ __forceinline__ bool in_faces_e(bool lo_on, bool hi_on, int a0, int a1, int n, int npml) {
    return (lo_on && meets(a0, a1, 2, npml)) || (hi_on && meets(a0, a1, n + 1 - npml, n - 1));
}

// TF/SF corrections of the H cells of a wavefront (row j, lanes k of z tile
// [kt0, kt1], planes i0..i1): Hy on the z faces k = k0-1 (+ dt/mu0
// den_hz(k) Ex_inc(k0)) and k = k1 (- ... Ex_inc(k1)), j in [j0, j1]; Hz on
// the y faces j = j0-1 (- dt/mu0 den_hy(j) Ex_inc(k)) and j = j1 (+), k in
// [k0, k1]; both for the planes i in [ia, ib] = [i0, i1-1] of the box.  The
// tile adds them before storing: H = (T)((double)H + c) with the lane's cy
// / cz (0 for lanes without a correction).  Returns false (uniform) when the
// wavefront has none.
template <typename T>
__device__ __forceinline__ bool tfsf_h_lane(const YeeArgs<T>& a, const TfsfLaunch& L, int i0, int i1, int j, int k,
                                            int kt0, int kt1, double& cy, double& cz, int& ia, int& ib) {
    const TfsfArgs& t = L.t;
    cy = 0.0;
    cz = 0.0;
    ia = max(i0, t.i0);
    ib = min(i1, t.i1 - 1);
    if (ia > ib) return false;
    const bool yrow = j >= t.j0 && j <= t.j1 && (meets(kt0, kt1, t.k0 - 1, t.k0 - 1) || meets(kt0, kt1, t.k1, t.k1));
    const bool zrow = (j == t.j0 - 1 || j == t.j1) && meets(kt0, kt1, t.k0, t.k1);
    if (!yrow && !zrow) return false;
    if (yrow && (k == t.k0 - 1 || k == t.k1)) {
        const bool hi = k == t.k1;
        const double c = t.dt_mu0 * (double)gld(a.den_hz, k) * L.ex1[(hi ? t.k1 : t.k0) - t.kb];
        cy = hi ? -c : c;
    }
    if (zrow && k >= t.k0 && k <= t.k1) {
        const double c = t.dt_mu0 * (double)gld(a.den_hy, j) * L.ex1[k - t.kb];
        cz = j == t.j1 ? c : -c;
    }
    return true;
}

// TF/SF corrections of the E cells of a wavefront (row j in [j0, j1], lanes
// k of z tile [kt0, kt1], planes i0..i1): Ex on the z faces k = k0 (+ dt/eps0
// den_ez(k) Hy_inc(k0-1)) and k = k1 (- ... Hy_inc(k1)) for the planes i in
// [xa, xb] = [i0, i1-1] of the box; Ez on the x faces i = i0 (cl = - dt/eps0
// den_ex(i0) Hy_inc(k)) and i = i1 (ch = + ...), k in [k0, k1-1].  The tile
// adds them before storing, as for H.  Returns false (uniform) when the
// wavefront has none.
template <typename T>
__device__ __forceinline__ bool tfsf_e_lane(const YeeArgs<T>& a, const TfsfLaunch& L, int i0, int i1, int j, int k,
                                            int kt0, int kt1, double& cx, double& cl, double& ch, int& xa, int& xb) {
    const TfsfArgs& t = L.t;
    cx = 0.0;
    cl = 0.0;
    ch = 0.0;
    xa = max(i0, t.i0);
    xb = min(i1, t.i1 - 1);
    if (j < t.j0 || j > t.j1) return false;
    const bool xrow = xa <= xb && (meets(kt0, kt1, t.k0, t.k0) || meets(kt0, kt1, t.k1, t.k1));
    const bool zrow = (meets(i0, i1, t.i0, t.i0) || meets(i0, i1, t.i1, t.i1)) && meets(kt0, kt1, t.k0, t.k1 - 1);
    if (!xrow && !zrow) return false;
    if (xrow && (k == t.k0 || k == t.k1)) {
        const bool hi = k == t.k1;
        const double c = t.dt_eps0 * (double)gld(a.den_ez, k) * (hi ? L.hy1[t.k1 - t.kb] : L.hy1[t.k0 - 1 - t.kb]);
        cx = hi ? -c : c;
    }
    if (!xrow) { xa = 1; xb = 0; }
    if (zrow && k >= t.k0 && k <= t.k1 - 1) {
        const double inc = L.hy1[k - t.kb];
        cl = -(t.dt_eps0 * (double)gld(a.den_ex, t.i0) * inc);
        ch = t.dt_eps0 * (double)gld(a.den_ex, t.i1) * inc;
    }
    return true;
}

// A tile: 64 lanes along z (one wavefront), one row along y per wavefront,
// marching xchunk planes along x.  A wavefront whose cells (its row, its z
// tile, the planes of its chunk) meet no CPML face runs the bulk loop.  In
// the face loop the face slots of the row (y) and lane (z) are fixed over
// the tile and those of the plane (x) are uniform; the x and y psi of the
// faces a cell lies in are loaded together with its fields, ahead of the
// arithmetic, and the z psi a batch of planes at a time (ZBatch).  A cell
// lies in at most one face per axis: grids are at least 2 npml cells along
// every CPML axis (omx3d_create rejects thinner ones).
template <typename T, bool CPML, bool B32, bool TF>
__device__ __forceinline__ void h_tile(const YeeArgs<T>& a, const TfsfLaunch& tf, const int t) {
    using O = Off<T, B32>;
    using I = typename O::I;
    const int tz = t % a.ntz;
    const int r = t / a.ntz;
    const int ty = r % a.nty;
    const int tx = r / a.nty;
    const int k = tz * kLanes + threadIdx.x + 1;
    const int j = ty * a.rows + __builtin_amdgcn_readfirstlane(threadIdx.y) + 1;
    const int i0 = tx * a.xchunk + 1;
    const int i1 = min(i0 + a.xchunk - 1, a.nx);
    const T* const cxs = sizeof(T) == 4 ? a.cdhx : a.den_hx;
    if (k > a.nz || j > a.ny) return;
    const I sx = O::of(a.sx), sz = O::of(a.sz), s1 = O::of(1);
    const long jke = (long)j * a.sz + k + a.off;
    I o = O::of((long)i0 * a.sx + jke);
    const T dhy = gld(a.den_hy, j), dhz = gld(a.den_hz, k);
    const T chy = sizeof(T) == 4 ? gld(a.cdhy, j) : (T)0, chz = sizeof(T) == 4 ? gld(a.cdhz, k) : (T)0;
    const T dmu = a.dt_mu0;
    T ezc = O::ld(a.ez, o), eyc = O::ld(a.ey, o);
    const int kt0 = tz * kLanes + 1, kt1 = min(kt0 + kLanes - 1, a.nz);
    double cy = 0.0, cz = 0.0;
    int ia = 0, ib = -1;
    const bool tfw = TF && tfsf_h_lane(a, tf, i0, i1, j, k, kt0, kt1, cy, cz, ia, ib);
    bool xyf = false, zf = false;
    if (CPML) {
        const unsigned cp = a.cpml;
        xyf = in_faces_h(cp & 1u, cp & 2u, i0, i1, a.nx, a.npml) || in_faces_h(cp & 4u, cp & 8u, j, j, a.ny, a.npml);
        zf = in_faces_h(cp & 16u, cp & 32u, kt0, kt1, a.nz, a.npml);
    }
    if (!xyf && !zf) {
        for (int i = i0; i <= i1; ++i, o += sx) {
            const T ezn = O::ld(a.ez, o + sx), eyn = O::ld(a.ey, o + sx);
            const T ezj = O::ld(a.ez, o + sz), eyk = O::ld(a.ey, o + s1);
            const T exc = O::ld(a.ex, o), exk = O::ld(a.ex, o + s1), exj = O::ld(a.ex, o + sz);
            T hx = O::ld(a.hx, o), hy = O::ld(a.hy, o), hz = O::ld(a.hz, o);
            const T cx = ldc(cxs, i);
            h_bulk(dmu, cx, dhy, dhz, chy, chz, ezc, ezn, ezj, eyc, eyn, eyk, exc, exk, exj, hx, hy, hz);
            if (tfw && i >= ia && i <= ib) { hy = (T)((double)hy + cy); hz = (T)((double)hz + cz); }
            O::st(a.hx, o, hx); O::st(a.hy, o, hy); O::st(a.hz, o, hz);
            ezc = ezn; eyc = eyn;
        }
        return;
    }
    // XY false: the wavefront meets only z faces (per lane), so the x and y
    // face code is compiled out of its loop
    const auto face_loop = [&](auto xy_const) {
        constexpr bool XY = decltype(xy_const)::value;
        const unsigned cp = a.cpml;
        const int npml = a.npml, nsh = 2 * (npml - 1);
        const Face fy = XY ? face_h(cp & 4u, cp & 8u, j, a.ny, npml) : Face{-1, 0};
        T yb = (T)0, yc = (T)0, zbc = (T)0, zcc = (T)0;
        if (fy.slot >= 0) { yb = gld(a.bh, fy.c); yc = gld(a.chd, fy.c); }
        // z faces: low k = 1..npml-1, high k = nz+1-npml..nz-1
        const ZBatch zw = z_batch((int)threadIdx.x, kt0, kt1, 1, (cp & 16u) ? npml - 1 : 0, a.nz + 1 - npml,
                                  (cp & 32u) ? a.nz - 1 : a.nz - npml, 1, j, i0, a.nz, a.sy, a.nx, npml);
        if (zw.q >= 0) { zbc = gld(a.bh, zw.m + 1); zcc = gld(a.chd, zw.m + 1); }
        const long zadv = (long)zw.P * (npml - 1);
        long zidx = zw.sidx;
        T zs0 = (T)0, zs1 = (T)0;
        int zb0 = i0;
        // psi of plane i: x (fx.slot*sx + jk), y (i*nsh*sz + slot*sz + k + off);
        // offsets of faces a cell is not in are unused
        const I jk = O::of(jke);
        const I ystep = O::of((long)nsh * a.sz);
        I py = O::of((long)i0 * nsh * a.sz + (long)fy.slot * a.sz + k + a.off);
        for (int i = i0; i <= i1; ++i, o += sx, py += ystep) {
            int zp = i - zb0;
            if (zw.nf > 0) {
                if (zp == zw.P) {
                    // the previous batch is complete: store it, start the next
                    if (zw.sp >= 0) { O::st(a.ph[4], O::of(zidx), zs0); O::st(a.ph[5], O::of(zidx), zs1); }
                    zb0 = i; zp = 0; zidx += zadv;
                }
                if (zp == 0 && zw.sp >= 0 && zb0 + zw.sp <= i1) {
                    zs0 = O::ld(a.ph[4], O::of(zidx)); zs1 = O::ld(a.ph[5], O::of(zidx));
                }
            }
            const T ezn = O::ld(a.ez, o + sx), eyn = O::ld(a.ey, o + sx);
            const T ezj = O::ld(a.ez, o + sz), eyk = O::ld(a.ey, o + s1);
            const T exc = O::ld(a.ex, o), exk = O::ld(a.ex, o + s1), exj = O::ld(a.ex, o + sz);
            T hx = O::ld(a.hx, o), hy = O::ld(a.hy, o), hz = O::ld(a.hz, o);
            const T cx = ldc(cxs, i);
            const Face fx = XY ? face_h(cp & 1u, cp & 2u, i, a.nx, npml) : Face{-1, 0};
            const I px = O::of((long)fx.slot * a.sx) + jk;
            T px0 = (T)0, px1 = (T)0, xb = (T)0, xc = (T)0, py0 = (T)0, py1 = (T)0;
            if (fx.slot >= 0) {
                px0 = O::ld(a.ph[0], px); px1 = O::ld(a.ph[1], px);
                xb = gld(a.bh, fx.c); xc = gld(a.chd, fx.c);
            }
            if (fy.slot >= 0) { py0 = O::ld(a.ph[2], py); py1 = O::ld(a.ph[3], py); }
            h_bulk(dmu, cx, dhy, dhz, chy, chz, ezc, ezn, ezj, eyc, eyn, eyk, exc, exk, exj, hx, hy, hz);
            if (fx.slot >= 0) {
                const T p0 = psi_step(px0, xb, xc, ezn - ezc), p1 = psi_step(px1, xb, xc, eyc - eyn);
                O::st(a.ph[0], px, p0); O::st(a.ph[1], px, p1);
                hy = hy + dmu * p0; hz = hz + dmu * p1;
            }
            if (fy.slot >= 0) {
                const T p0 = psi_step(py0, yb, yc, ezc - ezj), p1 = psi_step(py1, yb, yc, exj - exc);
                O::st(a.ph[2], py, p0); O::st(a.ph[3], py, p1);
                hx = hx + dmu * p0; hz = hz + dmu * p1;
            }
            if (zw.nf > 0) {
                // face lanes take their psi from the storage lanes and return
                // the new values to them (every active lane takes part)
                const int from = (zp * zw.nf + max(zw.q, 0)) * 4;
                const T p0 = psi_step(lane_pull(from, zs0), zbc, zcc, eyk - eyc);
                const T p1 = psi_step(lane_pull(from, zs1), zbc, zcc, exc - exk);
                if (zw.q >= 0) { hx = hx + dmu * p0; hy = hy + dmu * p1; }
                const T n0 = lane_pull(zw.slane * 4, p0), n1 = lane_pull(zw.slane * 4, p1);
                if (zw.sp == zp) { zs0 = n0; zs1 = n1; }
            }
            if (tfw && i >= ia && i <= ib) { hy = (T)((double)hy + cy); hz = (T)((double)hz + cz); }
            O::st(a.hx, o, hx); O::st(a.hy, o, hy); O::st(a.hz, o, hz);
            ezc = ezn; eyc = eyn;
        }
        if (zw.nf > 0 && zw.sp >= 0 && zb0 + zw.sp <= i1) {
            O::st(a.ph[4], O::of(zidx), zs0); O::st(a.ph[5], O::of(zidx), zs1);
        }
    };
    if (xyf) face_loop(std::true_type{});
    else face_loop(std::false_type{});
}

// ---------------------------------------------------------------------------
// E update: td_propagate_E_3Dfield (mxll_3D_mod.f90:742-921) with the per-cell
// medium update of get_medium_polarization (classical_medium_mod.f90:773-833).
// Currents of embedded quantum systems are applied by k_j_correct afterwards.
// ---------------------------------------------------------------------------
// p: the cell's position in the dispersive list, its index into the compact
// state eold, pd and into every pole plane of pl / plold (planes ps apart).
// The per-cell state of one component, loaded before the update.
template <typename T>
struct MedState {
    T eo, pv;                                        // E_old, PD
    T L[OMX_MAX_POLES], Lo[OMX_MAX_POLES];           // PL_k, PL_old_k
};

// A component's state loaded before its medium is known: every component's
// state block holds E_old, PD and maxp pole planes of PL and PL_old, so the
// loads depend on nothing but the cell (values a medium does not use are
// ignored); 0 for a vacuum component.
template <typename T>
__device__ __forceinline__ void medium_load_all(int m, long p, const T* __restrict__ st, long ps, int maxp,
                                                MedState<T>& s) {
    s.eo = (T)0; s.pv = (T)0;
#pragma unroll
    for (int q = 0; q < OMX_MAX_POLES; ++q) { s.L[q] = (T)0; s.Lo[q] = (T)0; }
    if (m == 0) return;
    s.eo = gld(st, p);
    s.pv = gld(st + ps, p);
#pragma unroll
    for (int q = 0; q < OMX_MAX_POLES; ++q)
        if (q < maxp) { s.L[q] = gld(st + (2 + q) * ps, p); s.Lo[q] = gld(st + (2 + maxp + q) * ps, p); }
}

// medium_update from a state loaded beforehand (medium_load_all): the same
// expressions in the same order (the fused step's post kernel, checked
// against the separate kernels' dispersive cells to 1e-13 in FP64), for a
// pass that loads every component's state before the first update.
template <typename T>
__device__ __forceinline__ T medium_apply(T e, T rot, T vac, int m, const MedDev<T>* __restrict__ med, long p,
                                          T* __restrict__ eold, T* __restrict__ pd, T* __restrict__ pl,
                                          T* __restrict__ plold, long ps, const MedState<T>& s) {
    if (m == 0) return e + vac;   // vac = dt_eps0 * rotH
    const OMX_GLOBAL MedDev<T>& M = *((const OMX_GLOBAL MedDev<T>*)med + m);
    if (M.kind == OMX_MEDIUM_DRUDE) {
        const T pv = s.pv;
        if constexpr (sizeof(T) == 8) {
            const T t = M.C1 * e + M.C3 * rot - M.C4 * pv;
            gst(pd, p, M.A1 * pv + M.A2 * (t + e));
            return t;
        } else {
            const T t = e + (M.C1 * e + M.C3 * rot - M.C4 * pv);
            gst(pd, p, pv + (M.A1 * pv + M.A2 * (t + e)));
            return t;
        }
    }
    if (M.kind == OMX_MEDIUM_DL) {
        const T eo = s.eo;
        const T pv = s.pv;
        const int np = M.np;
        T t;
        if constexpr (sizeof(T) == 8) {
            t = M.C1 * e + M.C2 * eo + M.C3 * rot - M.C4 * pv;
#pragma unroll
            for (int q = 0; q < OMX_MAX_POLES; ++q)
                if (q < np) t = t - M.B1[q] * s.L[q] - M.B2[q] * s.Lo[q];
            gst(pd, p, M.A1 * pv + M.A2 * (t + e));
        } else {
            T dlt = M.C1 * e + M.C2 * eo + M.C3 * rot - M.C4 * pv;
#pragma unroll
            for (int q = 0; q < OMX_MAX_POLES; ++q)
                if (q < np) dlt = dlt - M.B1[q] * s.L[q] - M.B2[q] * s.Lo[q];
            t = e + dlt;
            gst(pd, p, pv + (M.A1 * pv + M.A2 * (t + e)));
        }
        const T de = t - eo;
#pragma unroll
        for (int q = 0; q < OMX_MAX_POLES; ++q) {
            if (q < np) {
                const long pq = p + q * ps;
                const T L = s.L[q], Lo = s.Lo[q];
                T nl;
                if constexpr (sizeof(T) == 8) nl = M.al[q] * L + M.bt[q] * Lo + M.ga[q] * de;
                else nl = L + (((L - Lo) + M.al[q] * L) + M.bt[q] * Lo + M.ga[q] * de);
                gst(plold, pq, L);
                gst(pl, pq, nl);
            }
        }
        gst(eold, p, e);
        return t;
    }
    return e + M.cdt * rot;  // dielectric: dt_eps0 * (1/eps_r) * rotH
}

template <typename T>
__device__ __forceinline__ T medium_update(T e, T rot, T vac, int m, const MedDev<T>* __restrict__ med, long p,
                                           T* __restrict__ eold, T* __restrict__ pd, T* __restrict__ pl,
                                           T* __restrict__ plold, long ps) {
    if (m == 0) return e + vac;   // vac = dt_eps0 * rotH
    const OMX_GLOBAL MedDev<T>& M = *((const OMX_GLOBAL MedDev<T>*)med + m);
    if (M.kind == OMX_MEDIUM_DRUDE) {
        const T pv = gld(pd, p);
        if constexpr (sizeof(T) == 8) {
            const T t = M.C1 * e + M.C3 * rot - M.C4 * pv;
            gst(pd, p, M.A1 * pv + M.A2 * (t + e));
            return t;
        } else {
            const T t = e + (M.C1 * e + M.C3 * rot - M.C4 * pv);
            gst(pd, p, pv + (M.A1 * pv + M.A2 * (t + e)));
            return t;
        }
    }
    if (M.kind == OMX_MEDIUM_DL) {
        const T eo = gld(eold, p);
        const T pv = gld(pd, p);
        const int np = M.np;
        // the poles' PL / PL_old of the cell, loaded together (the loops are
        // unrolled to the maximum pole count, guarded by np) and kept in
        // registers for both passes: one memory round trip per component
        // instead of two per pole
        T Lq[OMX_MAX_POLES], Loq[OMX_MAX_POLES];
#pragma unroll
        for (int q = 0; q < OMX_MAX_POLES; ++q) {
            Lq[q] = (T)0; Loq[q] = (T)0;
            if (q < np) { Lq[q] = gld(pl, p + q * ps); Loq[q] = gld(plold, p + q * ps); }
        }
        T t;
        if constexpr (sizeof(T) == 8) {
            t = M.C1 * e + M.C2 * eo + M.C3 * rot - M.C4 * pv;
#pragma unroll
            for (int q = 0; q < OMX_MAX_POLES; ++q)
                if (q < np) t = t - M.B1[q] * Lq[q] - M.B2[q] * Loq[q];
            gst(pd, p, M.A1 * pv + M.A2 * (t + e));
        } else {
            T dlt = M.C1 * e + M.C2 * eo + M.C3 * rot - M.C4 * pv;
#pragma unroll
            for (int q = 0; q < OMX_MAX_POLES; ++q)
                if (q < np) dlt = dlt - M.B1[q] * Lq[q] - M.B2[q] * Loq[q];
            t = e + dlt;
            gst(pd, p, pv + (M.A1 * pv + M.A2 * (t + e)));
        }
        const T de = t - eo;
#pragma unroll
        for (int q = 0; q < OMX_MAX_POLES; ++q) {
            if (q < np) {
                const long pq = p + q * ps;
                const T L = Lq[q], Lo = Loq[q];
                T nl;
                if constexpr (sizeof(T) == 8) nl = M.al[q] * L + M.bt[q] * Lo + M.ga[q] * de;
                else nl = L + (((L - Lo) + M.al[q] * L) + M.bt[q] * Lo + M.ga[q] * de);
                gst(plold, pq, L);
                gst(pl, pq, nl);
            }
        }
        gst(eold, p, e);
        return t;
    }
    return e + M.cdt * rot;  // dielectric: dt_eps0 * (1/eps_r) * rotH
}

// The three components of a dispersive cell, each with its own medium (st0,
// st1, st2: the components' disjoint state blocks, E_old, PD, then the PL and
// PL_old planes, ps apart), each loading what its medium needs.
template <typename T>
__device__ __forceinline__ void medium_update3(T& ex, T& ey, T& ez, T rx, T ry, T rz, T vx, T vy, T vz, int m0, int m1,
                                               int m2, const MedDev<T>* __restrict__ med, long p,
                                               T* __restrict__ st0, T* __restrict__ st1, T* __restrict__ st2, long ps,
                                               int maxp) {
    const long pl0 = 2 * ps, plo0 = (2 + (long)maxp) * ps;
    ex = medium_update(ex, rx, vx, m0, med, p, st0, st0 + ps, st0 + pl0, st0 + plo0, ps);
    ey = medium_update(ey, ry, vy, m1, med, p, st1, st1 + ps, st1 + pl0, st1 + plo0, ps);
    ez = medium_update(ez, rz, vz, m2, med, p, st2, st2 + ps, st2 + pl0, st2 + plo0, ps);
}

// Curl of H at an E cell (rot, for the medium updates) and the vacuum
// increments dt_eps0 * rot (FP32: folded coefficients dt_eps0*den).
template <typename T>
__device__ __forceinline__ void e_curl(T dep, T hzc, T hzj, T hzi, T hyc, T hyk, T hyi, T hxc, T hxj, T hxk, T dex,
                                       T cex, T dey, T dez, T cey, T cez, T& rx, T& ry, T& rz, T& vx, T& vy, T& vz) {
    rx = (hzc - hzj) * dey + (hyk - hyc) * dez;
    ry = (hzi - hzc) * dex + (hxc - hxk) * dez;
    rz = (hyc - hyi) * dex + (hxj - hxc) * dey;
    if constexpr (sizeof(T) == 4) {
        vx = (hzc - hzj) * cey + (hyk - hyc) * cez;
        vy = (hzi - hzc) * cex + (hxc - hxk) * cez;
        vz = (hyc - hyi) * cex + (hxj - hxc) * cey;
    } else {
        vx = dep * rx; vy = dep * ry; vz = dep * rz;
    }
}

// Bulk E update of one cell in ex/ey/ez (no CPML).  tg is the cell's class
// tag (0: vacuum in all three components); dex = den_ex[i], cex_ = cdex[i]
// (FP32 only).  MED: 0 no media; 1 vacuum and dielectric components (the
// caller leaves cells of the Drude/Drude-Lorentz classes, tag >= disp_tag0,
// to the dispersive-cell pass).
template <typename T, int MED>
__device__ __forceinline__ void e_bulk(const YeeArgs<T>& a, int tg, T dep, T hzc, T hzj, T hzi, T hyc, T hyk, T hyi,
                                       T hxc, T hxj, T hxk, T dex, T cex_, T dey, T dez, T cey, T cez, T& ex, T& ey,
                                       T& ez) {
    static_assert(MED == 0 || MED == 1, "dispersive media are updated per cell by e_dispersive_cell");
    T rx, ry, rz, vx, vy, vz;
    e_curl(dep, hzc, hzj, hzi, hyc, hyk, hyi, hxc, hxj, hxk, dex, cex_, dey, dez, cey, cez, rx, ry, rz, vx, vy, vz);
    if (MED == 1 && tg != 0) {
        // vacuum: + dt_eps0 rotH; dielectric: + dt_eps0/eps_r rotH
        const int m0 = gld(a.cls, 3 * tg), m1 = gld(a.cls, 3 * tg + 1), m2 = gld(a.cls, 3 * tg + 2);
        ex = m0 == 0 ? ex + vx : ex + gld(&a.med[m0].cdt, 0) * rx;
        ey = m1 == 0 ? ey + vy : ey + gld(&a.med[m1].cdt, 0) * ry;
        ez = m2 == 0 ? ez + vz : ez + gld(&a.med[m2].cdt, 0) * rz;
    } else {
        ex = ex + vx; ey = ey + vy; ez = ez + vz;
    }
}

// x faces: psix_Eyx, psix_Ezx (low side i = 2..npml)
template <typename T, bool STEPPED = false>
__device__ __forceinline__ void e_cpml_x(const YeeArgs<T>& a, int i, long jk, T dep, T hzc, T hzi, T hyc, T hyi,
                                         T& ey, T& ez) {
    const unsigned cp = a.cpml;
    const int npml = a.npml;
    if ((cp & 1u) && i >= 2 && i <= npml) {
        const int c = i;
        const long ps = (long)(i - 1) * a.sx + jk;
        const T p0 = STEPPED ? gld(a.pe[0], ps) : psi_step(gld(a.pe[0], ps), gld(a.be, c), gld(a.ced, c), hzi - hzc);
        const T p1 = STEPPED ? gld(a.pe[1], ps) : psi_step(gld(a.pe[1], ps), gld(a.be, c), gld(a.ced, c), hyc - hyi);
        if (!STEPPED) { gst(a.pe[0], ps, p0); gst(a.pe[1], ps, p1); }
        ey = ey + dep * p0; ez = ez + dep * p1;
    }
    if ((cp & 2u) && i >= a.nx + 1 - npml && i <= a.nx - 1) {
        const int c = a.nx + 1 - i;
        const long ps = (long)(npml - 1 + c) * a.sx + jk;
        const T p0 = STEPPED ? gld(a.pe[0], ps) : psi_step(gld(a.pe[0], ps), gld(a.be, c), gld(a.ced, c), hzi - hzc);
        const T p1 = STEPPED ? gld(a.pe[1], ps) : psi_step(gld(a.pe[1], ps), gld(a.be, c), gld(a.ced, c), hyc - hyi);
        if (!STEPPED) { gst(a.pe[0], ps, p0); gst(a.pe[1], ps, p1); }
        ey = ey + dep * p0; ez = ez + dep * p1;
    }
}

// y faces: psiy_Exy, psiy_Ezy
template <typename T, bool STEPPED = false>
__device__ __forceinline__ void e_cpml_y(const YeeArgs<T>& a, int i, int j, int k, T dep, T hzc, T hzj, T hxc, T hxj,
                                         T& ex, T& ez) {
    const unsigned cp = a.cpml;
    const int npml = a.npml;
    const int nse = 2 * npml;
    if ((cp & 4u) && j >= 2 && j <= npml) {
        const int c = j;
        const long ps = ((long)i * nse + (j - 1)) * a.sz + k + a.off;
        const T p0 = STEPPED ? gld(a.pe[2], ps) : psi_step(gld(a.pe[2], ps), gld(a.be, c), gld(a.ced, c), hzc - hzj);
        const T p1 = STEPPED ? gld(a.pe[3], ps) : psi_step(gld(a.pe[3], ps), gld(a.be, c), gld(a.ced, c), hxj - hxc);
        if (!STEPPED) { gst(a.pe[2], ps, p0); gst(a.pe[3], ps, p1); }
        ex = ex + dep * p0; ez = ez + dep * p1;
    }
    if ((cp & 8u) && j >= a.ny + 1 - npml && j <= a.ny - 1) {
        const int c = a.ny + 1 - j;
        const long ps = ((long)i * nse + (npml - 1 + c)) * a.sz + k + a.off;
        const T p0 = STEPPED ? gld(a.pe[2], ps) : psi_step(gld(a.pe[2], ps), gld(a.be, c), gld(a.ced, c), hzc - hzj);
        const T p1 = STEPPED ? gld(a.pe[3], ps) : psi_step(gld(a.pe[3], ps), gld(a.be, c), gld(a.ced, c), hxj - hxc);
        if (!STEPPED) { gst(a.pe[2], ps, p0); gst(a.pe[3], ps, p1); }
        ex = ex + dep * p0; ez = ez + dep * p1;
    }
}

// z faces: psiz_Exz, psiz_Eyz
template <typename T, bool STEPPED = false>
__device__ __forceinline__ void e_cpml_z(const YeeArgs<T>& a, int i, int j, int k, T dep, T hyc, T hyk, T hxc, T hxk,
                                         T& ex, T& ey) {
    const unsigned cp = a.cpml;
    const int npml = a.npml;
    if ((cp & 16u) && k >= 2 && k <= npml) {
        const int c = k;
        const long ps = zpsi_index(0, j, i, k - 2, a.sy, a.nx, npml);
        const T p0 = STEPPED ? gld(a.pe[4], ps) : psi_step(gld(a.pe[4], ps), gld(a.be, c), gld(a.ced, c), hyk - hyc);
        const T p1 = STEPPED ? gld(a.pe[5], ps) : psi_step(gld(a.pe[5], ps), gld(a.be, c), gld(a.ced, c), hxc - hxk);
        if (!STEPPED) { gst(a.pe[4], ps, p0); gst(a.pe[5], ps, p1); }
        ex = ex + dep * p0; ey = ey + dep * p1;
    }
    if ((cp & 32u) && k >= a.nz + 1 - npml && k <= a.nz - 1) {
        const int c = a.nz + 1 - k;
        const long ps = zpsi_index(1, j, i, a.nz - 1 - k, a.sy, a.nx, npml);
        const T p0 = STEPPED ? gld(a.pe[4], ps) : psi_step(gld(a.pe[4], ps), gld(a.be, c), gld(a.ced, c), hyk - hyc);
        const T p1 = STEPPED ? gld(a.pe[5], ps) : psi_step(gld(a.pe[5], ps), gld(a.be, c), gld(a.ced, c), hxc - hxk);
        if (!STEPPED) { gst(a.pe[4], ps, p0); gst(a.pe[5], ps, p1); }
        ex = ex + dep * p0; ey = ey + dep * p1;
    }
}

// E tile (the tile and the bulk / face loops as in h_tile).  MED 0: no
// media; 1: vacuum and dielectric components, the cells of dispersive
// classes (tag >= disp_tag0) left entirely (E and psi of E) to
// k_e_dispersive.
template <typename T, bool CPML, int MED, bool B32, bool TF>
__device__ __forceinline__ void e_tile(const YeeArgs<T>& a, const TfsfLaunch& tf, const int t) {
    static_assert(MED == 0 || MED == 1, "the tiles take vacuum and dielectric cells");
    using O = Off<T, B32>;
    using I = typename O::I;
    const int tz = t % a.ntz;
    const int r = t / a.ntz;
    const int ty = r % a.nty;
    const int tx = r / a.nty;
    const int k = tz * kLanes + threadIdx.x + 1;
    const int j = ty * a.rows + __builtin_amdgcn_readfirstlane(threadIdx.y) + 1;
    const int i0 = tx * a.xchunk + 1;
    const int i1 = min(i0 + a.xchunk - 1, a.nx);
    if (k > a.nz || j > a.ny) return;
    const I sx = O::of(a.sx), sz = O::of(a.sz), s1 = O::of(1);
    const long jke = (long)j * a.sz + k + a.off;
    I o = O::of((long)i0 * a.sx + jke);
    const T dey = gld(a.den_ey, j), dez = gld(a.den_ez, k);
    const T cey = sizeof(T) == 4 ? gld(a.cdey, j) : (T)0, cez = sizeof(T) == 4 ? gld(a.cdez, k) : (T)0;
    const T dep = a.dt_eps0;
    const bool tags = MED != 0 && a.tag != nullptr;
    T hzi = O::ld(a.hz, o - sx), hyi = O::ld(a.hy, o - sx);
    const int kt0 = tz * kLanes + 1, kt1 = min(kt0 + kLanes - 1, a.nz);
    double tcx = 0.0, tcl = 0.0, tch = 0.0;
    int txa = 1, txb = 0;
    const bool tfw = TF && tfsf_e_lane(a, tf, i0, i1, j, k, kt0, kt1, tcx, tcl, tch, txa, txb);
    const int ti0 = TF ? tf.t.i0 : 0, ti1 = TF ? tf.t.i1 : 0;
    bool xyf = false, zf = false;
    if (CPML) {
        const unsigned cp = a.cpml;
        xyf = in_faces_e(cp & 1u, cp & 2u, i0, i1, a.nx, a.npml) || in_faces_e(cp & 4u, cp & 8u, j, j, a.ny, a.npml);
        zf = in_faces_e(cp & 16u, cp & 32u, kt0, kt1, a.nz, a.npml);
    }
    if (!xyf && !zf) {
        for (int i = i0; i <= i1; ++i, o += sx) {
            const T hzc = O::ld(a.hz, o), hyc = O::ld(a.hy, o), hxc = O::ld(a.hx, o);
            const T hzj = O::ld(a.hz, o - sz), hxj = O::ld(a.hx, o - sz);
            const T hyk = O::ld(a.hy, o - s1), hxk = O::ld(a.hx, o - s1);
            T ex = O::ld(a.ex, o), ey = O::ld(a.ey, o), ez = O::ld(a.ez, o);
            const int tg = tags ? O::tag(a.tag, o) : 0;
            const T dex = ldc(a.den_ex, i), cex = sizeof(T) == 4 ? ldc(a.cdex, i) : (T)0;
            if (MED != 1 || tg < a.disp_tag0) {
                e_bulk<T, MED>(a, tg, dep, hzc, hzj, hzi, hyc, hyk, hyi, hxc, hxj, hxk, dex, cex, dey, dez, cey,
                               cez, ex, ey, ez);
                if (tfw) {
                    if (i >= txa && i <= txb) ex = (T)((double)ex + tcx);
                    if (i == ti0) ez = (T)((double)ez + tcl);
                    if (i == ti1) ez = (T)((double)ez + tch);
                }
                O::st(a.ex, o, ex); O::st(a.ey, o, ey); O::st(a.ez, o, ez);
            }
            hzi = hzc; hyi = hyc;
        }
        return;
    }
    // XY false: the wavefront meets only z faces (per lane), so the x and y
    // face code is compiled out of its loop
    const auto face_loop = [&](auto xy_const) {
        constexpr bool XY = decltype(xy_const)::value;
        const unsigned cp = a.cpml;
        const int npml = a.npml, nse = 2 * npml;
        const Face fy = XY ? face_e(cp & 4u, cp & 8u, j, a.ny, npml) : Face{-1, 0};
        T yb = (T)0, yc = (T)0, zbc = (T)0, zcc = (T)0;
        if (fy.slot >= 0) { yb = gld(a.be, fy.c); yc = gld(a.ced, fy.c); }
        // z faces: low k = 2..npml, high k = nz+1-npml..nz-1
        const ZBatch zw = z_batch((int)threadIdx.x, kt0, kt1, 2, (cp & 16u) ? npml : 1, a.nz + 1 - npml,
                                  (cp & 32u) ? a.nz - 1 : a.nz - npml, 2, j, i0, a.nz, a.sy, a.nx, npml);
        if (zw.q >= 0) { zbc = gld(a.be, zw.m + 2); zcc = gld(a.ced, zw.m + 2); }
        const long zadv = (long)zw.P * (npml - 1);
        long zidx = zw.sidx;
        T zs0 = (T)0, zs1 = (T)0;
        int sown = 1;       // storage lane: its cell is not one of the dispersive workgroups'
        int zb0 = i0;
        const I jk = O::of(jke);
        const I ystep = O::of((long)nse * a.sz);
        I py = O::of((long)i0 * nse * a.sz + (long)fy.slot * a.sz + k + a.off);
        for (int i = i0; i <= i1; ++i, o += sx, py += ystep) {
            int zp = i - zb0;
            if (zw.nf > 0) {
                if (zp == zw.P) {
                    // the previous batch is complete: store it, start the next
                    if (zw.sp >= 0 && sown) { O::st(a.pe[4], O::of(zidx), zs0); O::st(a.pe[5], O::of(zidx), zs1); }
                    zb0 = i; zp = 0; zidx += zadv;
                }
                if (zp == 0 && zw.sp >= 0 && zb0 + zw.sp <= i1) {
                    zs0 = O::ld(a.pe[4], O::of(zidx)); zs1 = O::ld(a.pe[5], O::of(zidx));
                }
            }
            const T hzc = O::ld(a.hz, o), hyc = O::ld(a.hy, o), hxc = O::ld(a.hx, o);
            const T hzj = O::ld(a.hz, o - sz), hxj = O::ld(a.hx, o - sz);
            const T hyk = O::ld(a.hy, o - s1), hxk = O::ld(a.hx, o - s1);
            T ex = O::ld(a.ex, o), ey = O::ld(a.ey, o), ez = O::ld(a.ez, o);
            const int tg = tags ? O::tag(a.tag, o) : 0;
            const T dex = ldc(a.den_ex, i), cex = sizeof(T) == 4 ? ldc(a.cdex, i) : (T)0;
            const Face fx = XY ? face_e(cp & 1u, cp & 2u, i, a.nx, npml) : Face{-1, 0};
            const I px = O::of((long)fx.slot * a.sx) + jk;
            T px0 = (T)0, px1 = (T)0, xb = (T)0, xc = (T)0, py0 = (T)0, py1 = (T)0;
            if (fx.slot >= 0) {
                px0 = O::ld(a.pe[0], px); px1 = O::ld(a.pe[1], px);
                xb = gld(a.be, fx.c); xc = gld(a.ced, fx.c);
            }
            if (fy.slot >= 0) { py0 = O::ld(a.pe[2], py); py1 = O::ld(a.pe[3], py); }
            // cells of dispersive classes belong to the dispersive workgroups:
            // their loads above are unused and nothing of theirs is stored; the
            // psi loads are issued with the fields, ahead of the tag test
            const bool own = MED != 1 || tg < a.disp_tag0;
            e_bulk<T, MED>(a, tg, dep, hzc, hzj, hzi, hyc, hyk, hyi, hxc, hxj, hxk, dex, cex, dey, dez, cey, cez,
                           ex, ey, ez);
            if (fx.slot >= 0) {
                const T p0 = psi_step(px0, xb, xc, hzi - hzc), p1 = psi_step(px1, xb, xc, hyc - hyi);
                if (own) { O::st(a.pe[0], px, p0); O::st(a.pe[1], px, p1); }
                ey = ey + dep * p0; ez = ez + dep * p1;
            }
            if (fy.slot >= 0) {
                const T p0 = psi_step(py0, yb, yc, hzc - hzj), p1 = psi_step(py1, yb, yc, hxj - hxc);
                if (own) { O::st(a.pe[2], py, p0); O::st(a.pe[3], py, p1); }
                ex = ex + dep * p0; ez = ez + dep * p1;
            }
            if (zw.nf > 0) {
                // face lanes take their psi from the storage lanes and return
                // the new values (and whether they own the cell) to them
                const int from = (zp * zw.nf + max(zw.q, 0)) * 4;
                const T p0 = psi_step(lane_pull(from, zs0), zbc, zcc, hyk - hyc);
                const T p1 = psi_step(lane_pull(from, zs1), zbc, zcc, hxc - hxk);
                if (zw.q >= 0) { ex = ex + dep * p0; ey = ey + dep * p1; }
                const T n0 = lane_pull(zw.slane * 4, p0), n1 = lane_pull(zw.slane * 4, p1);
                const int no = __builtin_amdgcn_ds_bpermute(zw.slane * 4, own ? 1 : 0);
                if (zw.sp == zp) { zs0 = n0; zs1 = n1; sown = no; }
            }
            if (tfw) {
                if (i >= txa && i <= txb) ex = (T)((double)ex + tcx);
                if (i == ti0) ez = (T)((double)ez + tcl);
                if (i == ti1) ez = (T)((double)ez + tch);
            }
            if (own) { O::st(a.ex, o, ex); O::st(a.ey, o, ey); O::st(a.ez, o, ez); }
            hzi = hzc; hyi = hyc;
        }
        if (zw.nf > 0 && zw.sp >= 0 && zb0 + zw.sp <= i1 && sown) {
            O::st(a.pe[4], O::of(zidx), zs0); O::st(a.pe[5], O::of(zidx), zs1);
        }
    };
    if (xyf) face_loop(std::true_type{});
    else face_loop(std::false_type{});
}

// The E psi of a cell on CPML faces, loaded ahead of its update: the same
// faces, slots and coefficient indices as e_cpml_x / e_cpml_y / e_cpml_z (x,
// y, z: psi element, -1 off the faces; cx, cy, cz: coefficient index).
template <typename T>
struct PsiE {
    long x = -1, y = -1, z = -1;
    int cx = 0, cy = 0, cz = 0;
    T x0 = (T)0, x1 = (T)0, y0 = (T)0, y1 = (T)0, z0 = (T)0, z1 = (T)0;
};

template <typename T>
__device__ __forceinline__ void e_psi_load(const YeeArgs<T>& a, int i, int j, int k, long jk, PsiE<T>& q) {
    const unsigned cp = a.cpml;
    const int npml = a.npml, nse = 2 * npml;
    if ((cp & 1u) && i >= 2 && i <= npml) { q.cx = i; q.x = (long)(i - 1) * a.sx + jk; }
    else if ((cp & 2u) && i >= a.nx + 1 - npml && i <= a.nx - 1) {
        q.cx = a.nx + 1 - i; q.x = (long)(npml - 1 + q.cx) * a.sx + jk;
    }
    if ((cp & 4u) && j >= 2 && j <= npml) { q.cy = j; q.y = ((long)i * nse + (j - 1)) * a.sz + k + a.off; }
    else if ((cp & 8u) && j >= a.ny + 1 - npml && j <= a.ny - 1) {
        q.cy = a.ny + 1 - j; q.y = ((long)i * nse + (npml - 1 + q.cy)) * a.sz + k + a.off;
    }
    if ((cp & 16u) && k >= 2 && k <= npml) { q.cz = k; q.z = zpsi_index(0, j, i, k - 2, a.sy, a.nx, npml); }
    else if ((cp & 32u) && k >= a.nz + 1 - npml && k <= a.nz - 1) {
        q.cz = a.nz + 1 - k; q.z = zpsi_index(1, j, i, a.nz - 1 - k, a.sy, a.nx, npml);
    }
    if (q.x >= 0) { q.x0 = gld(a.pe[0], q.x); q.x1 = gld(a.pe[1], q.x); }
    if (q.y >= 0) { q.y0 = gld(a.pe[2], q.y); q.y1 = gld(a.pe[3], q.y); }
    if (q.z >= 0) { q.z0 = gld(a.pe[4], q.z); q.z1 = gld(a.pe[5], q.z); }
}

// e_cpml_x, e_cpml_y, e_cpml_z from the loaded psi (STEPPED: already
// advanced, added as they are).
template <typename T, bool STEPPED>
__device__ __forceinline__ void e_psi_apply(const YeeArgs<T>& a, const PsiE<T>& q, T dep, T hzc, T hzj, T hzi, T hyc,
                                            T hyk, T hyi, T hxc, T hxj, T hxk, T& ex, T& ey, T& ez) {
    if (q.x >= 0) {
        const T p0 = STEPPED ? q.x0 : psi_step(q.x0, gld(a.be, q.cx), gld(a.ced, q.cx), hzi - hzc);
        const T p1 = STEPPED ? q.x1 : psi_step(q.x1, gld(a.be, q.cx), gld(a.ced, q.cx), hyc - hyi);
        if (!STEPPED) { gst(a.pe[0], q.x, p0); gst(a.pe[1], q.x, p1); }
        ey = ey + dep * p0; ez = ez + dep * p1;
    }
    if (q.y >= 0) {
        const T p0 = STEPPED ? q.y0 : psi_step(q.y0, gld(a.be, q.cy), gld(a.ced, q.cy), hzc - hzj);
        const T p1 = STEPPED ? q.y1 : psi_step(q.y1, gld(a.be, q.cy), gld(a.ced, q.cy), hxj - hxc);
        if (!STEPPED) { gst(a.pe[2], q.y, p0); gst(a.pe[3], q.y, p1); }
        ex = ex + dep * p0; ez = ez + dep * p1;
    }
    if (q.z >= 0) {
        const T p0 = STEPPED ? q.z0 : psi_step(q.z0, gld(a.be, q.cz), gld(a.ced, q.cz), hyk - hyc);
        const T p1 = STEPPED ? q.z1 : psi_step(q.z1, gld(a.be, q.cz), gld(a.ced, q.cz), hxc - hxk);
        if (!STEPPED) { gst(a.pe[4], q.z, p0); gst(a.pe[5], q.z, p1); }
        ex = ex + dep * p0; ey = ey + dep * p1;
    }
}

// E of the cells of dispersive classes (Drude / Drude-Lorentz components),
// after the tile pass has updated every other cell: curl of H, the full
// medium update and the CPML psi recursions of the cell, from E in
// a.ex/a.ey/a.ez to the output arrays (the same arrays in the separate H/E
// path, the other buffer after the fused step).  list holds the cells'
// indices (i, j, k) with the media of their three components; p, the cell's
// position in its grid's list, indexes its compact Drude / DL state.
template <typename T, bool CPML, bool B32, bool STEPPED = false, bool LOADS_FIRST = false>
__device__ __forceinline__ void e_dispersive_cell(const YeeArgs<T>& a, T* __restrict__ exo, T* __restrict__ eyo,
                                                  T* __restrict__ ezo, long e, long p) {
    using O = Off<T, B32>;
    using I = typename O::I;
    // entry: i, j, k in bits 0-15, 16-31, 32-47; the media of the x, y, z
    // components in bits 48-51, 52-55, 56-59
    const int i = (int)(e & 0xffff), j = (int)((e >> 16) & 0xffff), k = (int)((e >> 32) & 0xffff);
    const int m0 = (int)((e >> 48) & 15), m1 = (int)((e >> 52) & 15), m2 = (int)((e >> 56) & 15);
    const long jk = (long)j * a.sz + k + a.off;
    const I o = O::of((long)i * a.sx + jk);
    const I sx = O::of(a.sx), sz = O::of(a.sz), s1 = O::of(1);
    // LOADS_FIRST: the three components' state and the E psi go out with the
    // fields, before any store (a latency-bound pass: a few memory round
    // trips per cell instead of one per component and face)
    MedState<T> ms0, ms1, ms2;
    PsiE<T> qe;
    if constexpr (LOADS_FIRST) {
        T* const st0 = a.dstate;
        medium_load_all(m0, p, st0, a.pstride, a.maxp, ms0);
        medium_load_all(m1, p, st0 + a.dcomp, a.pstride, a.maxp, ms1);
        medium_load_all(m2, p, st0 + 2 * a.dcomp, a.pstride, a.maxp, ms2);
        if constexpr (CPML) e_psi_load(a, i, j, k, jk, qe);
    }
    const T hzc = O::ld(a.hz, o), hyc = O::ld(a.hy, o), hxc = O::ld(a.hx, o);
    const T hzj = O::ld(a.hz, o - sz), hxj = O::ld(a.hx, o - sz);
    const T hyk = O::ld(a.hy, o - s1), hxk = O::ld(a.hx, o - s1);
    const T hzi = O::ld(a.hz, o - sx), hyi = O::ld(a.hy, o - sx);
    T ex = O::ld(a.ex, o), ey = O::ld(a.ey, o), ez = O::ld(a.ez, o);
    const T dex = gld(a.den_ex, i), cex = sizeof(T) == 4 ? gld(a.cdex, i) : (T)0;
    const T dey = gld(a.den_ey, j), dez = gld(a.den_ez, k);
    const T cey = sizeof(T) == 4 ? gld(a.cdey, j) : (T)0, cez = sizeof(T) == 4 ? gld(a.cdez, k) : (T)0;
    const T dep = a.dt_eps0;
    T rx, ry, rz, vx, vy, vz;
    e_curl(dep, hzc, hzj, hzi, hyc, hyk, hyi, hxc, hxj, hxk, dex, cex, dey, dez, cey, cez, rx, ry, rz, vx, vy, vz);
    if constexpr (LOADS_FIRST) {
        const long ps = a.pstride, pl0 = 2 * ps, plo0 = (2 + (long)a.maxp) * ps;
        T* const st0 = a.dstate;
        T* const st1 = st0 + a.dcomp;
        T* const st2 = st1 + a.dcomp;
        ex = medium_apply(ex, rx, vx, m0, a.med, p, st0, st0 + ps, st0 + pl0, st0 + plo0, ps, ms0);
        ey = medium_apply(ey, ry, vy, m1, a.med, p, st1, st1 + ps, st1 + pl0, st1 + plo0, ps, ms1);
        ez = medium_apply(ez, rz, vz, m2, a.med, p, st2, st2 + ps, st2 + pl0, st2 + plo0, ps, ms2);
        if constexpr (CPML)
            e_psi_apply<T, STEPPED>(a, qe, dep, hzc, hzj, hzi, hyc, hyk, hyi, hxc, hxj, hxk, ex, ey, ez);
    } else {
        medium_update3(ex, ey, ez, rx, ry, rz, vx, vy, vz, m0, m1, m2, a.med, p, a.dstate, a.dstate + a.dcomp,
                       a.dstate + 2 * a.dcomp, a.pstride, a.maxp);
        if constexpr (CPML) {
            // STEPPED: the fused step already advanced this cell's E psi (the
            // recursion depends on H only), so its values are added as they are
            e_cpml_x<T, STEPPED>(a, i, jk, dep, hzc, hzi, hyc, hyi, ey, ez);
            e_cpml_y<T, STEPPED>(a, i, j, k, dep, hzc, hzj, hxc, hxj, ex, ez);
            e_cpml_z<T, STEPPED>(a, i, j, k, dep, hyc, hyk, hxc, hxk, ex, ey);
        }
    }
    O::st(exo, o, ex); O::st(eyo, o, ey); O::st(ezo, o, ez);
}

template <typename T, bool CPML, bool B32>
__global__ __launch_bounds__(256) void k_e_dispersive(const YeeArgs<T> a, T* __restrict__ exo, T* __restrict__ eyo,
                                                      T* __restrict__ ezo, const long* __restrict__ list, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    e_dispersive_cell<T, CPML, B32>(a, exo, eyo, ezo, list[p], p);
}

// Minimum waves per SIMD the compiler must fit the E kernel into.  The
// FP32 E kernels on 32-bit offsets need a few more scalar registers than
// 8 waves leave; asked for 8, the compiler keeps the excess in vector-register
// lanes (no memory), and the launch runs 8 waves per SIMD instead of 7:
// E -3 % on the emitter slab, -5.5 % on the ellipsoids.  With TF/SF (and in
// the batched kernel) the step measured 0.5 % slower, and the FP64 and
// 64-bit-index kernels would spill to memory, so those keep 1 (no demand).
template <typename T, bool B32, bool TF>
inline constexpr int kEWaves = (sizeof(T) == 4 && B32 && !TF) ? 8 : 1;

// H tiles; with TF the tiles apply the TF/SF H corrections and the first
// tf.nb workgroups advance the 1D TF/SF line (hy1 from ex1) instead.
template <typename T, bool CPML, bool B32, bool TF>
__global__ __launch_bounds__(256) void k_h_update(const YeeArgs<T> a, const TfsfLaunch tf) {
    const int b = blockIdx.x;
    const int ns = TF ? tf.nb : 0;
    if (b < ns) {
        const int m = b * (int)(blockDim.x * blockDim.y) + (int)(threadIdx.y * blockDim.x + threadIdx.x);
        const TfsfArgs& t = tf.t;
        if (m < t.n1 - 1) tf.hy1[m] = tf.hy1[m] + t.dt_mu0 * t.inv_dr * (tf.ex1[m] - tf.ex1[m + 1]);
        return;
    }
    h_tile<T, CPML, B32, TF>(a, tf, tile_of_block(a.tile_order, b - ns, (int)gridDim.x - ns, a.ntz));
}

// E tiles; with MED 1 the first db workgroups take the cells of dispersive
// classes instead (one cell per thread, list as for k_e_dispersive): the
// tiles leave those cells alone, so both parts of the launch are
// independent, and the latency-bound per-cell work runs beside the
// streaming tiles.
// With TF the tiles add the TF/SF E corrections and workgroups
// [db, db + tf.nb) advance the 1D TF/SF line's E (ex1 from hy1, plus the
// source at node ks); the tiles read only hy1.
template <typename T, bool CPML, int MED, bool B32, bool TF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kEWaves<T, B32, TF>, 8)))
void k_e_update(const YeeArgs<T> a, const long* __restrict__ disp, int n_disp, int db, const TfsfLaunch tf) {
    const int b = blockIdx.x;
    if (MED == 1 && b < db) {
        const int p = (b * (int)blockDim.y + (int)threadIdx.y) * kLanes + (int)threadIdx.x;
        if (p < n_disp) e_dispersive_cell<T, CPML, B32>(a, a.ex, a.ey, a.ez, disp[p], p);
        return;
    }
    const int ns = db + (TF ? tf.nb : 0);
    if (b < ns) {
        const int m = (b - db) * (int)(blockDim.x * blockDim.y) + (int)(threadIdx.y * blockDim.x + threadIdx.x);
        const TfsfArgs& t = tf.t;
        if (m >= 1 && m < t.n1 - 1) {
            double v = tf.ex1[m] + t.dt_eps0 * t.inv_dr * (tf.hy1[m - 1] - tf.hy1[m]);
            if (m == t.ks - t.kb) v += tf.src;
            tf.ex1[m] = v;
        }
        return;
    }
    // the E sweep runs backwards through the tiles: it starts where the H
    // sweep ended, so the fields H wrote last (and read last) are still in
    // the memory-side cache, and it ends where the next H sweep starts
    const int nt = (int)gridDim.x - ns;
    e_tile<T, CPML, MED, B32, TF>(a, tf, nt - 1 - tile_of_block(a.tile_order, b - ns, nt, a.ntz));
}

// ---------------------------------------------------------------------------
// Batched launches: one launch covers the tiles of many independent grids.
// tile_start[g] is the first tile of grid g (tile_start[n] = total); each
// workgroup finds its grid by binary search and runs the single-grid tile
// code on that grid's argument block, so every grid's arithmetic is
// identical to its own single-grid launch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_grid(const int* __restrict__ tile_start, int n, int t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Tile of launch block b of a batch: the grids' blocks follow one another,
// each grid's range padded to a multiple of 8 (tile_start: the ranges'
// starts, n+1 entries), so that the blocks of one XCD keep one residue of 8
// inside every grid and each grid deals its tiles in its own order
// (a.tile_order over its own z tiles).  -1: a padding block.
template <typename T>
__device__ __forceinline__ int batch_tile(const YeeArgs<T>* __restrict__ args, const int* __restrict__ tile_start,
                                          int n, int b, int& g) {
    g = find_grid(tile_start, n, b);
    const YeeArgs<T>& a = args[g];
    const int nb = tile_start[g + 1] - tile_start[g];
    const int t = tile_of_block(a.tile_order, b - tile_start[g], nb, a.ntz);
    const int nt = a.ntz * a.nty * ((a.nx + a.xchunk - 1) / a.xchunk);
    return t < nt ? t : -1;
}

template <typename T, bool B32>
__global__ __launch_bounds__(256) void k_h_update_batched(const YeeArgs<T>* __restrict__ args,
                                                          const int* __restrict__ tile_start, int n) {
    int g;
    const int t = batch_tile(args, tile_start, n, (int)blockIdx.x, g);
    if (t < 0) return;
    const YeeArgs<T>& a = args[g];
    if ((int)threadIdx.y >= a.rows) return;
    const TfsfLaunch none{};
    if (a.cpml) h_tile<T, true, B32, false>(a, none, t);
    else h_tile<T, false, B32, false>(a, none, t);
}

// Batched E tiles; the first db workgroups take the dispersive cells of
// every grid (disp: the grids' lists one after the other, each padded with
// -1 to whole workgroups; grid_of[b]: the grid of workgroup b).
template <typename T, bool B32>
__global__ __launch_bounds__(256) void k_e_update_batched(const YeeArgs<T>* __restrict__ args,
                                                          const int* __restrict__ tile_start, int n,
                                                          const long* __restrict__ disp,
                                                          const int* __restrict__ grid_of, int n_disp, int db) {
    const int b = blockIdx.x;
    if (b < db) {
        const int p = (b * (int)blockDim.y + (int)threadIdx.y) * kLanes + (int)threadIdx.x;
        const long e = p < n_disp ? disp[p] : -1L;
        if (e < 0) return;
        // the CPML terms test each face's bit of a.cpml, so one instance
        // serves grids with and without CPML
        const YeeArgs<T>& a = args[grid_of[b]];
        e_dispersive_cell<T, true, B32>(a, a.ex, a.ey, a.ez, e, p - a.disp0);
        return;
    }
    // backwards, as in k_e_update: the last grid's last tiles first
    int g;
    const int t = batch_tile(args, tile_start, n, tile_start[n] - 1 - (b - db), g);
    if (t < 0) return;
    const YeeArgs<T>& a = args[g];
    if ((int)threadIdx.y >= a.rows) return;
    const TfsfLaunch none{};
    if (a.cpml) e_tile<T, true, 1, B32, false>(a, none, t);
    else e_tile<T, false, 1, B32, false>(a, none, t);
}

// ---------------------------------------------------------------------------
// Periodic wrap: whole-plane copies before each half step, non-MPI branch of
// mxll_3D_mod.f90:528-541 (E) and :725-738 (H).  One launch per axis, in the
// reference's x, y, z order.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_wrap(T* __restrict__ f0, T* __restrict__ f1, Layout L, int axis, int dst, int src) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    int na, nb;
    if (axis == 0) { na = L.nz + 2; nb = L.ny + 2; }
    else if (axis == 1) { na = L.nz + 2; nb = L.nx + 2; }
    else { na = L.ny + 2; nb = L.nx + 2; }
    if (u >= na || v >= nb) return;
    long od, os;
    if (axis == 0) { od = L.at(dst, v, u); os = L.at(src, v, u); }
    else if (axis == 1) { od = L.at(v, dst, u); os = L.at(v, src, u); }
    else { od = L.at(v, u, dst); os = L.at(v, u, src); }
    f0[od] = f0[os];
    f1[od] = f1[os];
}

// ---------------------------------------------------------------------------
// Soft current source: update_source (source_mod.f90:106-132) and the 3D
// branch of source_interactions (interactions_mod.f90:102-156).  The J_mat
// entry one step past the kernel edge reads as zero.
// ---------------------------------------------------------------------------
struct SourceParams {
    double j_amp, w0, t0, tau, t_init, t_final, phase, c_src, dt;
};

// J_mat = j_amp * exp(-((t-t0)/tau)^2) * cos(w0 (t-t0) + phase) * ker, zero
// outside [t_init, t_final]; E += c_src * 1/2 (J_mat(p) + J_mat(p+1)).
__device__ inline double source_amp(const SourceParams& s, double time) {
    if (!(time >= s.t_init && time <= s.t_final)) return 0.0;
    const double x = (time - s.t0) / s.tau;
    return (s.j_amp * exp(-(x * x))) * cos(s.w0 * (time - s.t0) + s.phase);
}

// E += add at one field point; a point inside a near-field patch (sh != null)
// is accumulated in its FP64 shadow and the field gets the rounded shadow.
template <typename T>
__device__ __forceinline__ void add_to_e(T* f, long o, double* sh, double add) {
    if (sh) {
        const double v = *sh + add;
        *sh = v;
        f[o] = (T)v;
    } else {
        f[o] = (T)((double)f[o] + add);
    }
}

// sh: per point, its FP64 shadow element or null (sh itself may be null).
template <typename T>
__global__ void k_source(T* __restrict__ f, const long* __restrict__ off, const double* __restrict__ ka,
                         const double* __restrict__ kb, int n, SourceParams sp, double time,
                         double* const* __restrict__ sh) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double s = source_amp(sp, time);
    const double jav = 0.5 * (s * ka[p] + s * kb[p]);
    add_to_e(f, off[p], sh ? sh[p] : nullptr, sp.c_src * jav);
}

// All points of all sources of a batch; the step index comes from device
// memory so that the launch can be replayed from a captured graph.  Points
// of one source never share a cell; sources that share a cell are kept in
// separate launches (pass), in their order.
template <typename T>
__global__ void k_source_batched(T* const* __restrict__ field, const long* __restrict__ off,
                                 const double* __restrict__ ka, const double* __restrict__ kb,
                                 const int* __restrict__ src_of, const SourceParams* __restrict__ sp, int n,
                                 const long* __restrict__ step, double* const* __restrict__ sh) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const SourceParams& q = sp[src_of[p]];
    const double time = q.dt * ((double)step[0] - 0.5);
    const double s = source_amp(q, time);
    const double jav = 0.5 * (s * ka[p] + s * kb[p]);
    add_to_e(field[p], off[p], sh[p], q.c_src * jav);
}

__global__ void k_step_advance(long* step) { step[0] += 1; }

// ---------------------------------------------------------------------------
// Detector sampling with the stencils of write_3D_field
// (write_fields_subs_mod.f90:813-1509); point detectors read Ez unaveraged.
// code = 2*field + (point ? 1 : 0).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_sample(const YeeArgs<T> a, const int* __restrict__ code, const long* __restrict__ off,
                         long n, double* __restrict__ out) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int cd = code[s];
    const int field = cd >> 1;
    const bool point = cd & 1;
    const long o = off[s];
    const long sx = a.sx, sz = a.sz;
    double v = 0.0;
    switch (field) {
        case OMX_EX: v = 0.5 * ((double)a.ex[o - sx] + (double)a.ex[o]); break;
        case OMX_EY: v = 0.5 * ((double)a.ey[o - sz] + (double)a.ey[o]); break;
        case OMX_EZ: v = point ? (double)a.ez[o] : 0.5 * ((double)a.ez[o - 1] + (double)a.ez[o]); break;
        case OMX_HX: v = 0.25 * ((((double)a.hx[o - sz] + (double)a.hx[o]) + (double)a.hx[o - 1]) + (double)a.hx[o - sz - 1]); break;
        case OMX_HY: v = 0.25 * ((((double)a.hy[o - sx] + (double)a.hy[o]) + (double)a.hy[o - 1]) + (double)a.hy[o - sx - 1]); break;
        case OMX_HZ: v = 0.25 * ((((double)a.hz[o] + (double)a.hz[o - sx]) + (double)a.hz[o - sz]) + (double)a.hz[o - sx - sz]); break;
        default: break;
    }
    out[s] = v;
}

struct BatchRecord {
    double* rec;        // record buffer of the grid
    long n_samples;     // samples per row of the grid
    int every, row0, cap;
};

// Detector samples of all grids of a batch at step `step`, when
// mod(step, every) == 0; row = step/every - row0 of that grid.
template <typename T>
__global__ void k_sample_batched(const YeeArgs<T>* __restrict__ args, const BatchRecord* __restrict__ br,
                                 const int* __restrict__ grid_of, const int* __restrict__ local,
                                 const int* __restrict__ code, const long* __restrict__ off, long n,
                                 const long* __restrict__ step, int* __restrict__ overflow) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int g = grid_of[s];
    const BatchRecord& R = br[g];
    const long st = step[0];
    if (R.every <= 0 || st % R.every != 0) return;
    const long row = st / R.every - R.row0;
    if (row < 0 || row >= R.cap) { overflow[0] = 1; return; }
    const YeeArgs<T>& a = args[g];
    const int cd = code[s];
    const int field = cd >> 1;
    const bool point = cd & 1;
    const long o = off[s];
    const long sx = a.sx, sz = a.sz;
    double v = 0.0;
    switch (field) {
        case OMX_EX: v = 0.5 * ((double)a.ex[o - sx] + (double)a.ex[o]); break;
        case OMX_EY: v = 0.5 * ((double)a.ey[o - sz] + (double)a.ey[o]); break;
        case OMX_EZ: v = point ? (double)a.ez[o] : 0.5 * ((double)a.ez[o - 1] + (double)a.ez[o]); break;
        case OMX_HX: v = 0.25 * ((((double)a.hx[o - sz] + (double)a.hx[o]) + (double)a.hx[o - 1]) + (double)a.hx[o - sz - 1]); break;
        case OMX_HY: v = 0.25 * ((((double)a.hy[o - sx] + (double)a.hy[o]) + (double)a.hy[o - 1]) + (double)a.hy[o - sx - 1]); break;
        case OMX_HZ: v = 0.25 * ((((double)a.hz[o] + (double)a.hz[o - sx]) + (double)a.hz[o - sz]) + (double)a.hz[o - sx - sz]); break;
        default: break;
    }
    R.rec[row * R.n_samples + local[s]] = v;
}

// ---------------------------------------------------------------------------
// Embedded quantum systems ("material" groups).
// ---------------------------------------------------------------------------
// E at each site: send_E_3D_to_J_ranks, non-MPI (interactions_mod.f90:656-691).
template <typename T>
__global__ void k_gather_e(const YeeArgs<T> a, const long* __restrict__ off, int n, double* __restrict__ e) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long o = off[q];
    e[3 * q + 0] = 0.5 * ((double)a.ex[o - a.sx] + (double)a.ex[o]);
    e[3 * q + 1] = 0.5 * ((double)a.ey[o - a.sz] + (double)a.ey[o]);
    e[3 * q + 2] = 0.5 * ((double)a.ez[o - 1] + (double)a.ez[o]);
}

// J_old <- J; J(cell) = coef * dPt/dt; dJ = (J - J_old)/dt_q on the cells
// the q-groups write, group after group (send_J_to_E_3D_ranks,
// interactions_mod.f90:1067-1069,1179-1210,1216-1218): the last writer of a
// cell wins, as in the reference's sequential loops, and with several
// groups only the last group's change survives in dJ.
__global__ void k_scatter_j(const QSlot* __restrict__ slots, int nslot, const double* __restrict__ dpdt,
                            double dt_q, double* __restrict__ J) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const QSlot q = slots[s];
    for (int c = 0; c < 3; ++c) {
        const double nw = q.coef * dpdt[3 * q.sys + c];
        if (q.mode == 1) {
            J[9 * s + 3 + c] = nw;
            J[9 * s + c] = nw;
            J[9 * s + 6 + c] = 0.0;
        } else {
            const double old = q.prev >= 0 ? q.coef_prev * dpdt[3 * q.prev + c] : J[9 * s + c];
            J[9 * s + 3 + c] = old;
            J[9 * s + c] = nw;
            J[9 * s + 6 + c] = (nw - old) / dt_q;
        }
    }
}

// E -= dt_eps0 * Jav at no-medium cells next to a site (the J term of
// mxll_3D_mod.f90:757-760,780-783,805-808).  a: slot of cell i, b: slot of
// cell i+1 along the component axis (-1: none, J = 0 there).
template <typename T>
__global__ void k_j_correct(T* __restrict__ f, const long* __restrict__ off, const int* __restrict__ sa,
                            const int* __restrict__ sb, int n, int comp, const double* __restrict__ J,
                            double tau, double dt_eps0, const uint8_t* __restrict__ tag,
                            const int8_t* __restrict__ cls, const int* __restrict__ kind,
                            double* const* __restrict__ sh) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long o = off[q];
    if (tag) {
        const int m = cls[3 * (int)tag[o] + comp];
        if (m != 0 && (kind[m] == OMX_MEDIUM_DRUDE || kind[m] == OMX_MEDIUM_DL)) return;
    }
    const int A = sa[q], B = sb[q];
    const double ja = A >= 0 ? J[9 * A + 3 + comp] : 0.0;
    const double da = A >= 0 ? J[9 * A + 6 + comp] : 0.0;
    const double jb = B >= 0 ? J[9 * B + 3 + comp] : 0.0;
    const double db = B >= 0 ? J[9 * B + 6 + comp] : 0.0;
    const double jav = 0.5 * (((ja + tau * da) + jb) + tau * db);
    add_to_e(f, o, sh ? sh[q] : nullptr, -(dt_eps0 * jav));
}

// Two-level emitter, FP64 RK4 over one quantum step with the gathered field
// held constant.  State per site (stride 16):
//   0 omega0  1 mu  2..4 unit polarisation  5 gamma2 (dephasing rate)
//   6 rho_gg  7 rho_ee  8 Re rho_eg  9 Im rho_eg  10..12 dipole  13..15 dip_old
// H = diag(-w0/2, w0/2) - mu E_p sigma_x (hbar = 1); coherences decay at gamma2.
// dipole = 2 mu Re(rho_eg) * pol; dPt/dt = (dipole - dip_old)/dt_q.
__device__ inline void tls_rhs(double w0, double V, double g2, const double y[4], double d[4]) {
    const double pop = y[0] - y[1];
    d[0] = 2.0 * V * y[3];
    d[1] = -2.0 * V * y[3];
    d[2] = w0 * y[3] - g2 * y[2];
    d[3] = -w0 * y[2] - V * pop - g2 * y[3];
}

__global__ void k_tls_step(double* __restrict__ st, const double* __restrict__ e, int n, double dt_q,
                           double* __restrict__ dpdt) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double* s = st + 16 * q;
    const double w0 = s[0], mu = s[1], g2 = s[5];
    const double ep = e[3 * q] * s[2] + e[3 * q + 1] * s[3] + e[3 * q + 2] * s[4];
    const double V = -mu * ep;
    double y[4] = {s[6], s[7], s[8], s[9]};
    double k1[4], k2[4], k3[4], k4[4], yt[4];
    const double h = dt_q;
    tls_rhs(w0, V, g2, y, k1);
    for (int c = 0; c < 4; ++c) yt[c] = y[c] + 0.5 * h * k1[c];
    tls_rhs(w0, V, g2, yt, k2);
    for (int c = 0; c < 4; ++c) yt[c] = y[c] + 0.5 * h * k2[c];
    tls_rhs(w0, V, g2, yt, k3);
    for (int c = 0; c < 4; ++c) yt[c] = y[c] + h * k3[c];
    tls_rhs(w0, V, g2, yt, k4);
    for (int c = 0; c < 4; ++c) y[c] = y[c] + (h / 6.0) * (k1[c] + 2.0 * k2[c] + 2.0 * k3[c] + k4[c]);
    s[6] = y[0]; s[7] = y[1]; s[8] = y[2]; s[9] = y[3];
    const double amp = 2.0 * mu * y[2];
    for (int c = 0; c < 3; ++c) {
        const double dn = amp * s[2 + c];
        const double dold = s[10 + c];
        s[13 + c] = dold;
        s[10 + c] = dn;
        dpdt[3 * q + c] = (dn - dold) / dt_q;
    }
}

// ---------------------------------------------------------------------------
// Energy: W = 1/2 dr^3 sum (eps0 eps_r E^2 + mu0 H^2), FP64 and compensated
// FP32 (Neumaier) partial sums per workgroup, then one final workgroup.
// ---------------------------------------------------------------------------
__device__ inline void neumaier(float& s, float& c, float x) {
    const float t = s + x;
    if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

template <typename T, bool MEDIA>
__global__ __launch_bounds__(256) void k_energy(const YeeArgs<T> a, const double* __restrict__ eps_cls,
                                                double* __restrict__ part) {
    __shared__ double sd[256];
    __shared__ float ss[256], sc[256];
    const long ncell = (long)a.nx * a.ny * a.nz;
    double acc = 0.0;
    float fs = 0.0f, fc = 0.0f;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
        const int k = (int)(c % a.nz) + 1;
        const int j = (int)((c / a.nz) % a.ny) + 1;
        const int i = (int)(c / ((long)a.nz * a.ny)) + 1;
        const long o = (long)i * a.sx + (long)j * a.sz + k + a.off;
        double ex = a.ex[o], ey = a.ey[o], ez = a.ez[o];
        double hx = a.hx[o], hy = a.hy[o], hz = a.hz[o];
        double er[3] = {1.0, 1.0, 1.0};
        if (MEDIA && a.tag != nullptr) {
            const int tg = a.tag[o];
            er[0] = eps_cls[3 * tg]; er[1] = eps_cls[3 * tg + 1]; er[2] = eps_cls[3 * tg + 2];
        }
        const double we = kEps0 * ((er[0] * ex * ex + er[1] * ey * ey) + er[2] * ez * ez);
        const double wh = kMu0 * ((hx * hx + hy * hy) + hz * hz);
        acc += we + wh;
        const float fex = (float)a.ex[o], fey = (float)a.ey[o], fez = (float)a.ez[o];
        const float fhx = (float)a.hx[o], fhy = (float)a.hy[o], fhz = (float)a.hz[o];
        neumaier(fs, fc, (float)kEps0 * ((float)er[0] * fex * fex));
        neumaier(fs, fc, (float)kEps0 * ((float)er[1] * fey * fey));
        neumaier(fs, fc, (float)kEps0 * ((float)er[2] * fez * fez));
        neumaier(fs, fc, (float)kMu0 * (fhx * fhx));
        neumaier(fs, fc, (float)kMu0 * (fhy * fhy));
        neumaier(fs, fc, (float)kMu0 * (fhz * fhz));
    }
    const int t = threadIdx.x;
    sd[t] = acc; ss[t] = fs; sc[t] = fc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            sd[t] += sd[t + w];
            float s = ss[t], cc = sc[t] + sc[t + w];
            neumaier(s, cc, ss[t + w]);
            ss[t] = s; sc[t] = cc;
        }
        __syncthreads();
    }
    if (t == 0) {
        part[3 * blockIdx.x] = sd[0];
        part[3 * blockIdx.x + 1] = (double)ss[0];
        part[3 * blockIdx.x + 2] = (double)sc[0];
    }
}

__global__ __launch_bounds__(256) void k_energy_final(const double* __restrict__ part, int nb, double scale,
                                                      double* __restrict__ out) {
    __shared__ double sd[256];
    __shared__ float ss[256], sc[256];
    const int t = threadIdx.x;
    double acc = 0.0;
    float fs = 0.0f, fc = 0.0f;
    for (int b = t; b < nb; b += 256) {
        acc += part[3 * b];
        neumaier(fs, fc, (float)part[3 * b + 1]);
        fc += (float)part[3 * b + 2];
    }
    sd[t] = acc; ss[t] = fs; sc[t] = fc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            sd[t] += sd[t + w];
            float s = ss[t], cc = sc[t] + sc[t + w];
            neumaier(s, cc, ss[t + w]);
            ss[t] = s; sc[t] = cc;
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = 0.5 * scale * sd[0];
        out[1] = 0.5 * scale * (double)(ss[0] + sc[0]);
    }
}

// ---------------------------------------------------------------------------
// Near-field patches of FP32 grids.  Next to a point source the field is
// dominated by its quasi-static near field, some (c/(omega dr))^2 times the
// radiated field at one cell (1e3-1e4 at the workloads' resolutions);
// rounding that field to FP32 every half step injects noise that radiates
// to the detectors at ~1e-4 of their signal.  A box of cells around each
// source keeps FP64 shadows of E and H: after the FP32 H (E) launch the box's
// H (E) is computed again in FP64 from the shadows inside the box and the
// FP32 fields outside it, and both the shadow and the FP32 field get the
// result.  Boxes lie outside the CPML and away from the outer planes, and
// two boxes never touch (the host merges them), so a neighbour of a box cell
// is either in the same box or in no box.  Cells of Drude / Drude-Lorentz
// classes in a box keep their ADE state in FP64 as well.  The arithmetic is
// that of the FP64 engine (h_bulk, e_curl, e_bulk, medium_update).
// ---------------------------------------------------------------------------
// (PatchBox: omx_internal.hpp)

__device__ __forceinline__ int find_box(const PatchBox* __restrict__ b, int n, long q) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b[mid].first <= q) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Field value at (i,j,k): the shadow when the cell is in the box, else the
// FP32 field (neighbours of box cells are in the box or in no box).
template <typename T>
__device__ __forceinline__ double patch_val(const PatchBox& B, const double* sh, const T* f, const YeeArgs<T>& a,
                                            int i, int j, int k) {
    const int di = i - B.lo[0], dj = j - B.lo[1], dk = k - B.lo[2];
    if ((unsigned)di < (unsigned)B.w[0] && (unsigned)dj < (unsigned)B.w[1] && (unsigned)dk < (unsigned)B.w[2])
        return sh[((long)di * B.w[1] + dj) * B.w[2] + dk];
    return (double)f[(long)i * a.sx + (long)j * a.sz + k + a.off];
}

// Cell (i,j,k) of box cell r (k fastest); 32-bit divisions by the box's
// uniform extents.
__device__ __forceinline__ void patch_cell(const PatchBox& B, int r, int& i, int& j, int& k) {
    const unsigned w2 = (unsigned)B.w[2], w1 = (unsigned)B.w[1];
    const unsigned t = (unsigned)r / w2;
    const unsigned u = t / w1;
    k = B.lo[2] + (int)((unsigned)r - t * w2);
    j = B.lo[1] + (int)(t - u * w1);
    i = B.lo[0] + (int)u;
}

// H of a box cell from the E of a.ex/a.ey/a.ez (E^n) into the shadow and
// a.hx/a.hy/a.hz.
template <typename T>
__device__ __forceinline__ void patch_h_cell(const YeeArgs<T>& a, const PatchBox& B, int r) {
    int i, j, k;
    patch_cell(B, r, i, j, k);
    const double* ex = B.pe;
    const double* ey = B.pe + B.cs;
    const double* ez = B.pe + 2 * B.cs;
    const double ezc = patch_val(B, ez, a.ez, a, i, j, k), ezj = patch_val(B, ez, a.ez, a, i, j + 1, k);
    const double ezn = patch_val(B, ez, a.ez, a, i + 1, j, k);
    const double eyc = patch_val(B, ey, a.ey, a, i, j, k), eyk = patch_val(B, ey, a.ey, a, i, j, k + 1);
    const double eyn = patch_val(B, ey, a.ey, a, i + 1, j, k);
    const double exc = patch_val(B, ex, a.ex, a, i, j, k), exk = patch_val(B, ex, a.ex, a, i, j, k + 1);
    const double exj = patch_val(B, ex, a.ex, a, i, j + 1, k);
    const double dhx = B.den[3][i], dhy = B.den[4][j], dhz = B.den[5][k], dmu = B.dt_mu0;
    const double hx = B.ph[r] + dmu * ((ezc - ezj) * dhy + (eyk - eyc) * dhz);
    const double hy = B.ph[B.cs + r] + dmu * ((ezn - ezc) * dhx + (exc - exk) * dhz);
    const double hz = B.ph[2 * B.cs + r] + dmu * ((eyc - eyn) * dhx + (exj - exc) * dhy);
    B.ph[r] = hx; B.ph[B.cs + r] = hy; B.ph[2 * B.cs + r] = hz;
    const long o = (long)i * a.sx + (long)j * a.sz + k + a.off;
    a.hx[o] = (T)hx; a.hy[o] = (T)hy; a.hz[o] = (T)hz;
}

// One E component of a box cell of a dispersive class: the update of
// medium_update (get_medium_polarization, classical_medium_mod.f90:773-833)
// in FP64 on the box's state (c: component, m: its medium, r: box cell, dp:
// the cell's position in the dispersive list); the FP32 compact state of the
// cell takes the rounded values.  vac_c: the vacuum / dielectric coefficient.
template <typename T>
__device__ __forceinline__ double patch_medium(const YeeArgs<T>& a, const PatchBox& B, int c, int m, long r, int dp,
                                               double e, double rot, double vac_c) {
    if (m == 0) return e + vac_c * rot;
    const int kind = a.med[m].kind;
    if (kind != OMX_MEDIUM_DRUDE && kind != OMX_MEDIUM_DL) return e + vac_c * rot;
    const omx_medium_coeffs& M = B.mco[m];
    const long cs = B.cs;
    double* const s0 = B.ds + (long)c * B.nv * cs + r;     // E_old, PD, PL_q, PL_old_q at s0 + v * cs
    T* const f0 = a.dstate + c * a.dcomp + dp;             // the same in the FP32 compact state
    const long ps = a.pstride;
    const int maxp = a.maxp;
    const double pv = s0[cs];
    double t;
    if (kind == OMX_MEDIUM_DRUDE) {
        t = M.C1 * e + M.C3 * rot - M.C4 * pv;
        const double pd = M.A1 * pv + M.A2 * (t + e);
        s0[cs] = pd;
        f0[ps] = (T)pd;
        return t;
    }
    const int np = a.med[m].np;
    const double eo = s0[0];
    t = M.C1 * e + M.C2 * eo + M.C3 * rot - M.C4 * pv;
    for (int q = 0; q < np; ++q) t = t - M.B1_k[q] * s0[(2 + q) * cs] - M.B2_k[q] * s0[(2 + maxp + q) * cs];
    const double pd = M.A1 * pv + M.A2 * (t + e);
    s0[cs] = pd;
    f0[ps] = (T)pd;
    const double de = t - eo;
    for (int q = 0; q < np; ++q) {
        const double L = s0[(2 + q) * cs], Lo = s0[(2 + maxp + q) * cs];
        const double nl = M.alpha_k[q] * L + M.beta_k[q] * Lo + M.gamma_k[q] * de;
        s0[(2 + maxp + q) * cs] = L;
        s0[(2 + q) * cs] = nl;
        f0[(2 + maxp + q) * ps] = (T)L;
        f0[(2 + q) * ps] = (T)nl;
    }
    s0[0] = e;
    f0[0] = (T)e;
    return t;
}

// E of a box cell from the H of a.hx/a.hy/a.hz (H^{n+1/2}) into the shadow
// and a.ex/a.ey/a.ez; a cell of a dispersive class keeps the FP32 result of
// the dispersive update and its shadow takes that value.
template <typename T>
__device__ __forceinline__ void patch_e_cell(const YeeArgs<T>& a, const PatchBox& B, int r) {
    int i, j, k;
    patch_cell(B, r, i, j, k);
    const long o = (long)i * a.sx + (long)j * a.sz + k + a.off;
    const int tg = a.tag ? (int)a.tag[o] : 0;
    const int dp = tg >= a.disp_tag0 && B.dl ? B.dl[r] : -1;
    if (tg >= a.disp_tag0 && dp < 0) {
        B.pe[r] = (double)a.ex[o]; B.pe[B.cs + r] = (double)a.ey[o]; B.pe[2 * B.cs + r] = (double)a.ez[o];
        return;
    }
    const double* hx = B.ph;
    const double* hy = B.ph + B.cs;
    const double* hz = B.ph + 2 * B.cs;
    const double hzc = patch_val(B, hz, a.hz, a, i, j, k), hzj = patch_val(B, hz, a.hz, a, i, j - 1, k);
    const double hzi = patch_val(B, hz, a.hz, a, i - 1, j, k);
    const double hyc = patch_val(B, hy, a.hy, a, i, j, k), hyk = patch_val(B, hy, a.hy, a, i, j, k - 1);
    const double hyi = patch_val(B, hy, a.hy, a, i - 1, j, k);
    const double hxc = patch_val(B, hx, a.hx, a, i, j, k), hxk = patch_val(B, hx, a.hx, a, i, j, k - 1);
    const double hxj = patch_val(B, hx, a.hx, a, i, j - 1, k);
    const double dex = B.den[0][i], dey = B.den[1][j], dez = B.den[2][k];
    const double rx = (hzc - hzj) * dey + (hyk - hyc) * dez;
    const double ry = (hzi - hzc) * dex + (hxc - hxk) * dez;
    const double rz = (hyc - hyi) * dex + (hxj - hxc) * dey;
    double ex, ey, ez;
    if (dp < 0) {
        ex = B.pe[r] + B.cdt[3 * tg] * rx;
        ey = B.pe[B.cs + r] + B.cdt[3 * tg + 1] * ry;
        ez = B.pe[2 * B.cs + r] + B.cdt[3 * tg + 2] * rz;
    } else {
        // a cell of a dispersive class: the full per-cell medium update in
        // FP64 on the box's state, which the FP32 state then takes rounded
        const int m0 = a.cls[3 * tg], m1 = a.cls[3 * tg + 1], m2 = a.cls[3 * tg + 2];
        ex = patch_medium(a, B, 0, m0, r, dp, B.pe[r], rx, B.cdt[3 * tg]);
        ey = patch_medium(a, B, 1, m1, r, dp, B.pe[B.cs + r], ry, B.cdt[3 * tg + 1]);
        ez = patch_medium(a, B, 2, m2, r, dp, B.pe[2 * B.cs + r], rz, B.cdt[3 * tg + 2]);
    }
    B.pe[r] = ex; B.pe[B.cs + r] = ey; B.pe[2 * B.cs + r] = ez;
    a.ex[o] = (T)ex; a.ey[o] = (T)ey; a.ez[o] = (T)ez;
}

// Shadows from the FP32 fields (set-up, and after fields were written from
// the host).
template <typename T>
__device__ __forceinline__ void patch_sync_cell(const YeeArgs<T>& a, const PatchBox& B, int r) {
    int i, j, k;
    patch_cell(B, r, i, j, k);
    const long o = (long)i * a.sx + (long)j * a.sz + k + a.off;
    B.pe[r] = (double)a.ex[o]; B.pe[B.cs + r] = (double)a.ey[o]; B.pe[2 * B.cs + r] = (double)a.ez[o];
    B.ph[r] = (double)a.hx[o]; B.ph[B.cs + r] = (double)a.hy[o]; B.ph[2 * B.cs + r] = (double)a.hz[o];
    const int dp = B.dl ? B.dl[r] : -1;
    if (dp >= 0)
        for (int c = 0; c < 3; ++c)
            for (int v = 0; v < B.nv; ++v)
                B.ds[((long)c * B.nv + v) * B.cs + r] = (double)a.dstate[c * a.dcomp + (long)v * a.pstride + dp];
}

// phase 0: H, 1: E, 2: shadows from the fields.  Every box starts on a
// multiple of the workgroup size (256) in the launch's flat index, so a
// workgroup's box is found once, from uniform values; the padding threads
// past a box's cells return.  One grid: the argument block by value; ncell:
// the padded cells of all boxes.
template <typename T>
__global__ __launch_bounds__(256) void k_patch(const YeeArgs<T> a, const PatchBox* __restrict__ boxes, int nbox,
                                               long ncell, int phase) {
    const long q0 = (long)blockIdx.x * blockDim.x;
    if (q0 >= ncell) return;
    const PatchBox& B = boxes[find_box(boxes, nbox, q0)];
    const int r = (int)(q0 - B.first) + (int)threadIdx.x;
    if (r >= B.w[0] * B.w[1] * B.w[2]) return;
    if (phase == 0) patch_h_cell(a, B, r);
    else if (phase == 1) patch_e_cell(a, B, r);
    else patch_sync_cell(a, B, r);
}

// The boxes of every grid of a batch: box B belongs to grid B.g.
template <typename T>
__global__ __launch_bounds__(256) void k_patch_batched(const YeeArgs<T>* __restrict__ args,
                                                       const PatchBox* __restrict__ boxes, int nbox, long ncell,
                                                       int phase) {
    const long q0 = (long)blockIdx.x * blockDim.x;
    if (q0 >= ncell) return;
    const PatchBox& B = boxes[find_box(boxes, nbox, q0)];
    const int r = (int)(q0 - B.first) + (int)threadIdx.x;
    if (r >= B.w[0] * B.w[1] * B.w[2]) return;
    const YeeArgs<T>& a = args[B.g];
    if (phase == 0) patch_h_cell(a, B, r);
    else if (phase == 1) patch_e_cell(a, B, r);
    else patch_sync_cell(a, B, r);
}

// ---------------------------------------------------------------------------
// TF/SF plane wave (TfsfArgs above): the E corrections and the 1D line's E
// half step after the E launch or the fused step (the separate H and E
// launches otherwise carry the corrections and the line's half steps
// themselves, and the fused step the H corrections and the line's H half
// step).
// ---------------------------------------------------------------------------

// Item q of the TF/SF E step (ex, ez: the arrays holding E^{n+1}): Ex on the
// box's z faces (q < 2 sa), Ez on its x faces (the next 2 sb), then the 1D
// line's E update (n1 items).
template <typename T>
__device__ __forceinline__ void tfsf_e_item(const YeeArgs<T>& a, T* __restrict__ ex, T* __restrict__ ez,
                                            const TfsfArgs& t, double* __restrict__ ex1,
                                            const double* __restrict__ hy1, double src, int q) {
    const int nxi = t.i1 - t.i0, nyj = t.j1 - t.j0 + 1;
    const int nyz = t.j1 - t.j0 + 1, nzk = t.k1 - t.k0;
    const int sa = nxi * nyj, sb = nyz * nzk;
    if (q < 2 * sa) {                        // Ex on the z faces
        const int side = q / sa, r = q % sa;
        const int i = t.i0 + r % nxi, j = t.j0 + r / nxi;
        if (side == 0) {
            const int k = t.k0;
            const long o = a.sx * i + (long)a.sz * j + k + a.off;
            ex[o] = (T)((double)ex[o] + t.dt_eps0 * (double)a.den_ez[k] * hy1[t.k0 - 1 - t.kb]);
        } else {
            const int k = t.k1;
            const long o = a.sx * i + (long)a.sz * j + k + a.off;
            ex[o] = (T)((double)ex[o] - t.dt_eps0 * (double)a.den_ez[k] * hy1[t.k1 - t.kb]);
        }
        return;
    }
    const int q2 = q - 2 * sa;
    if (q2 < 2 * sb) {                       // Ez on the x faces (consecutive threads along z)
        const int side = q2 / sb, r = q2 % sb;
        const int k = t.k0 + r % nzk, j = t.j0 + r / nzk;
        const double inc = hy1[k - t.kb];
        if (side == 0) {
            const int i = t.i0;
            const long o = a.sx * i + (long)a.sz * j + k + a.off;
            ez[o] = (T)((double)ez[o] - t.dt_eps0 * (double)a.den_ex[i] * inc);
        } else {
            const int i = t.i1;
            const long o = a.sx * i + (long)a.sz * j + k + a.off;
            ez[o] = (T)((double)ez[o] + t.dt_eps0 * (double)a.den_ex[i] * inc);
        }
        return;
    }
    const int m = q2 - 2 * sb;               // 1D E update (ends held at zero)
    if (m >= 1 && m < t.n1 - 1) {
        double v = ex1[m] + t.dt_eps0 * t.inv_dr * (hy1[m - 1] - hy1[m]);
        if (m == t.ks - t.kb) v += src;
        ex1[m] = v;
    }
}

template <typename T>
__global__ void k_tfsf_e(const YeeArgs<T> a, TfsfArgs t, double* __restrict__ ex1, const double* __restrict__ hy1,
                         double src) {
    tfsf_e_item<T>(a, a.ex, a.ez, t, ex1, hy1, src, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// Sets *flag when a listed dispersive cell (entries as for
// e_dispersive_cell) lies on the E correction faces of the TF/SF box.
__global__ void k_disp_on_tfsf(const long* __restrict__ list, int n, TfsfArgs t, int* __restrict__ flag) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long e = list[p];
    const int i = (int)(e & 0xffff), j = (int)((e >> 16) & 0xffff), k = (int)((e >> 32) & 0xffff);
    const bool ex_face = (k == t.k0 || k == t.k1) && i >= t.i0 && i <= t.i1 - 1 && j >= t.j0 && j <= t.j1;
    const bool ez_face = (i == t.i0 || i == t.i1) && j >= t.j0 && j <= t.j1 && k >= t.k0 && k <= t.k1 - 1;
    if (ex_face || ez_face) flag[0] = 1;
}

}  // namespace omx

#include "omx_fused.hpp"
