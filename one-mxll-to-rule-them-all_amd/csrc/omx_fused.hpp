// omx_fused.hpp -- the fused H+E step (step mode 0) of libomx_fdtd_hip.
//
// One pass per time step: E^n and H^{n-1/2} are read from one buffer set and
// E^{n+1} and H^{n+1/2} written to the other, so each field array is read
// once and written once per step (12 words per cell instead of the 18 of
// separate H and E passes).  The arithmetic of every cell is that of the
// separate kernels (h_tile / e_tile: td_propagate_H_3Dfield and
// td_propagate_E_3Dfield, mxll_3D_mod.f90:505-923), face order included.
//
// A wavefront owns 64 cells along z (lanes), R rows along y and a chunk of
// planes along x, which it marches carrying the next plane's E and the
// previous plane's H in registers.  The E update of a cell needs H^{n+1/2}
// at j-1, k-1 and i-1: the rows of the wavefront take j-1 from the row
// before (the first row from a halo row computed again), lanes take k-1 from
// the lane before by a DPP shift (lane 0 computes its k0-1 cell itself),
// planes take i-1 from the plane before (the chunk's first plane from a halo
// plane computed again).  Halo cells run the same H routine as their owners
// with the CPML psi of H read from the input buffer and nothing stored, so
// their values equal the owner's bit for bit; the psi of H is therefore
// double-buffered like the fields.  E(k+1) comes by the opposite DPP shift,
// lane 63 loading its k0+64 neighbour.  No LDS and no barrier.
//
// Cells of Drude / Drude-Lorentz classes are left to the dispersive pass
// after the kernel (their E and E psi are not stored); the TF/SF H
// corrections are applied inside (owned and halo cells alike), the E side
// after the kernel.
#pragma once

namespace omx {

__device__ __forceinline__ float dpp_shr1(float old, float v) {   // lane l <- lane l-1; lane 0 <- old
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_shl1(float old, float v) {   // lane l <- lane l+1; lane 63 <- old
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ double dpp_shr1(double old, double v) {
    const unsigned long long o = __builtin_bit_cast(unsigned long long, old), w = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)w, 0x138, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(w >> 32), 0x138,
                                                              0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double dpp_shl1(double old, double v) {
    const unsigned long long o = __builtin_bit_cast(unsigned long long, old), w = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)w, 0x130, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(w >> 32), 0x130,
                                                              0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Outputs and layout of the fused step (the inputs are the YeeArgs block).
template <typename T>
struct FusedOut {
    T *ex, *ey, *ez, *hx, *hy, *hz;   // E^{n+1}, H^{n+1/2}
    T* ph[6];                         // psi of H at n+1/2 (x, y, z slab arrays as YeeArgs::ph)
    const double* ex1;                // TF/SF: the incident line's E^n (null without TF/SF)
    TfsfArgs t;
    int tile_order;                   // 0: block order, 1: one contiguous slab of tiles per XCD
    // The launch covers up to four boxes of tiles, numbered one box after
    // the other: planes i_lo..i_hi, rows j_lo..j_hi, z tiles tz_lo ..
    // tz_lo+ntz-1 (z tile fastest, then row block of R rows, then chunk of
    // xc planes); t_end is one past the box's last tile.
    struct Box {
        int i_lo, i_hi, j_lo, j_hi, tz_lo, ntz, nty, ntx, xc, t_end;
    } box[4];
    int nbox;
};

// Per-row, per-lane or per-plane coefficients of the H and E updates and
// the CPML face there: den_h, dt_mu0*den_h (FP32), den_e, dt_eps0*den_e
// (FP32); sh / se locate the psi of the H / E face (-1: no face): the slot
// of an x or y face, or, for a lane in a z face, the element index of its
// psi for (j, i) = (0, 0) (the psi of cell (i,j,k) is then
// sh + (j*(nx+2) + i)*(npml-1)); bh, chh, be, cee are the face's psi
// coefficients (b-1 in FP32).
template <typename T>
struct Coef {
    T dh, ch, de, ce;
    int sh, se;
    T bh, chh, be, cee;
};

// n: the index along the axis (1-based); Z: the z axis (sh / se in the
// z-face form); U: n is uniform over the wavefront (scalar loads), else
// per lane.
template <typename T, bool CPML, bool Z, bool U>
__device__ __forceinline__ Coef<T> coef_at(const YeeArgs<T>& a, const T* den_h, const T* cd_h, const T* den_e,
                                           const T* cd_e, int n, int na, unsigned lo_bit, unsigned hi_bit) {
    Coef<T> c;
    c.dh = U ? ldc(den_h, n) : gld(den_h, n);
    c.de = U ? ldc(den_e, n) : gld(den_e, n);
    c.ch = sizeof(T) == 4 ? (U ? ldc(cd_h, n) : gld(cd_h, n)) : (T)0;
    c.ce = sizeof(T) == 4 ? (U ? ldc(cd_e, n) : gld(cd_e, n)) : (T)0;
    c.sh = -1; c.se = -1;
    c.bh = c.chh = c.be = c.cee = (T)0;
    if (CPML) {
        const Face fh = face_h(a.cpml & lo_bit, a.cpml & hi_bit, n, na, a.npml);
        const Face fe = face_e(a.cpml & lo_bit, a.cpml & hi_bit, n, na, a.npml);
        if (fh.slot >= 0) {
            c.bh = U ? ldc(a.bh, fh.c) : gld(a.bh, fh.c);
            c.chh = U ? ldc(a.chd, fh.c) : gld(a.chd, fh.c);
            // z face: side 0 for k <= npml-1, cell m = c-1 of its side
            c.sh = Z ? (int)zpsi_index(n <= a.npml - 1 ? 0 : 1, 0, 0, fh.c - 1, a.sy, a.nx, a.npml) : fh.slot;
        }
        if (fe.slot >= 0) {
            c.be = U ? ldc(a.be, fe.c) : gld(a.be, fe.c);
            c.cee = U ? ldc(a.ced, fe.c) : gld(a.ced, fe.c);
            // z face: side 0 for k <= npml, cell m = c-2 of its side
            c.se = Z ? (int)zpsi_index(n <= a.npml ? 0 : 1, 0, 0, fe.c - 2, a.sy, a.nx, a.npml) : fe.slot;
        }
    }
    return c;
}

// TF/SF corrections of H in z cell k (same values as tfsf_h_lane), read
// once per wavefront: cy is the Hy correction of the cells of the box's z
// faces (k = k0-1 or k1) for rows j0..j1, e1 the incident Ex at k (k0..k1,
// else 0) that the Hz correction of rows j0-1 and j1 scales.  The cell's
// plane decides whether either applies.
struct TfLane {
    double cy, e1;
};
template <typename T, bool U>
__device__ __forceinline__ TfLane tf_lane(const YeeArgs<T>& a, const FusedOut<T>& fo, int k) {
    const TfsfArgs& t = fo.t;
    TfLane v{0.0, 0.0};
    if (k == t.k0 - 1 || k == t.k1) {
        const bool hi = k == t.k1;
        const double c = t.dt_mu0 * (double)(U ? ldc(a.den_hz, k) : gld(a.den_hz, k)) * fo.ex1[(hi ? t.k1 : t.k0) - t.kb];
        v.cy = hi ? -c : c;
    }
    if (k >= t.k0 && k <= t.k1) v.e1 = fo.ex1[k - t.kb];
    return v;
}

// CPML psi offsets of cell (i,j,k) (bytes; x and y faces: slot form, z
// faces: the lane's z-face base) and the psi values read ahead of their use
// (zero where the cell has no such face).  Reading a row's psi together
// with its fields keeps them in one memory round trip instead of one per
// face.
struct PsiOff {
    uint32_t x, y, z;
};
template <typename T>
struct Psi {
    T x0, x1, y0, y1, z0, z1;   // x face (Hy,Hz | Ey,Ez), y face (Hx,Hz | Ex,Ez), z face (Hx,Hy | Ex,Ey)
};

// H (E false) or E (E true) psi offsets of cell (i,j,k).
template <typename T, bool E>
__device__ __forceinline__ PsiOff psi_off(const YeeArgs<T>& a, int i, int j, int k, const Coef<T>& px,
                                          const Coef<T>& ry, const Coef<T>& lz) {
    const uint32_t S = (uint32_t)sizeof(T);
    const uint32_t kb = (uint32_t)(k + a.off);
    const int sx = E ? px.se : px.sh, sy = E ? ry.se : ry.sh, sz = E ? lz.se : lz.sh;
    const uint32_t ny2 = (uint32_t)(E ? 2 * a.npml : 2 * (a.npml - 1));
    PsiOff o;
    o.x = ((uint32_t)sx * (uint32_t)a.sx + (uint32_t)j * (uint32_t)a.sz + kb) * S;
    o.y = (((uint32_t)i * ny2 + (uint32_t)sy) * (uint32_t)a.sz + kb) * S;
    o.z = ((uint32_t)sz + ((uint32_t)j * (uint32_t)(a.nx + 2) + (uint32_t)i) * (uint32_t)(a.npml - 1)) * S;
    return o;
}

// Reads the psi of H (E false: the components WX/WY/WZ need) or of E (E
// true: all) of the faces cell (i,j,k) lies in.
template <typename T, int CPML, bool E, bool WX, bool WY, bool WZ>
__device__ __forceinline__ Psi<T> psi_read(const YeeArgs<T>& a, const PsiOff& o, const Coef<T>& px, const Coef<T>& ry,
                                           const Coef<T>& lz) {
    using O = Off<T, true>;
    Psi<T> v;
    v.x0 = v.x1 = v.y0 = v.y1 = v.z0 = v.z1 = (T)0;
    if constexpr (CPML != 0) {
        T* const* ps = E ? a.pe : a.ph;
        if ((CPML & 1) && (E ? px.se : px.sh) >= 0) {
            if (E || WY) v.x0 = O::ld(ps[0], o.x);
            if (E || WZ) v.x1 = O::ld(ps[1], o.x);
        }
        if ((CPML & 2) && (E ? ry.se : ry.sh) >= 0) {
            if (E || WX) v.y0 = O::ld(ps[2], o.y);
            if (E || WZ) v.y1 = O::ld(ps[3], o.y);
        }
        if ((CPML & 4) && (E ? lz.se : lz.sh) >= 0) {
            if (E || WX) v.z0 = O::ld(ps[4], o.z);
            if (E || WY) v.z1 = O::ld(ps[5], o.z);
        }
    }
    return v;
}

// H^{n+1/2} of the components WX/WY/WZ of cell (i,j,k) (h in: H^{n-1/2}):
// the bulk curl, then the CPML psi of the faces the cell lies in (x, y, z,
// as the reference; ps: psi of H at n-1/2 read by psi_read), then the TF/SF
// correction (tl: the cell's z cell's values, tf_lane).  STORE writes the new psi to fo.ph at po where act is set
// (owned cells).
// cx is the plane's coefficient (cdhx[i] in FP32, den_hx[i] in FP64) and px
// its face; ry / lz the row's and the lane's coefficients and faces.
// CPML: bit 0 x faces, bit 1 y faces, bit 2 z faces.
template <typename T, int CPML, bool TF, bool WX, bool WY, bool WZ, bool STORE>
__device__ __forceinline__ void h_fused(const YeeArgs<T>& a, const FusedOut<T>& fo, int i, int j, int k, T cx,
                                        const Coef<T>& px, const Coef<T>& ry, const Coef<T>& lz, const Psi<T>& ps,
                                        const PsiOff& po, const TfLane& tl, T ezc, T ezj, T ezn, T eyc, T eyk, T eyn,
                                        T exc, T exk, T exj, T& hx, T& hy, T& hz, bool act = true) {
    using O = Off<T, true>;
    const T dmu = a.dt_mu0;
    if constexpr (sizeof(T) == 4) {
        if constexpr (WX) hx = hx + (ezc - ezj) * ry.ch + (eyk - eyc) * lz.ch;
        if constexpr (WY) hy = hy + (ezn - ezc) * cx + (exc - exk) * lz.ch;
        if constexpr (WZ) hz = hz + (eyc - eyn) * cx + (exj - exc) * ry.ch;
    } else {
        if constexpr (WX) hx = hx + dmu * ((ezc - ezj) * ry.dh + (eyk - eyc) * lz.dh);
        if constexpr (WY) hy = hy + dmu * ((ezn - ezc) * cx + (exc - exk) * lz.dh);
        if constexpr (WZ) hz = hz + dmu * ((eyc - eyn) * cx + (exj - exc) * ry.dh);
    }
    if constexpr (CPML != 0) {
        if ((CPML & 1) && px.sh >= 0 && (WY || WZ)) {
            if constexpr (WY) {
                const T p0 = psi_step(ps.x0, px.bh, px.chh, ezn - ezc);
                if (STORE && act) O::st(fo.ph[0], po.x, p0);
                hy = hy + dmu * p0;
            }
            if constexpr (WZ) {
                const T p1 = psi_step(ps.x1, px.bh, px.chh, eyc - eyn);
                if (STORE && act) O::st(fo.ph[1], po.x, p1);
                hz = hz + dmu * p1;
            }
        }
        if ((CPML & 2) && ry.sh >= 0 && (WX || WZ)) {
            if constexpr (WX) {
                const T p0 = psi_step(ps.y0, ry.bh, ry.chh, ezc - ezj);
                if (STORE && act) O::st(fo.ph[2], po.y, p0);
                hx = hx + dmu * p0;
            }
            if constexpr (WZ) {
                const T p1 = psi_step(ps.y1, ry.bh, ry.chh, exj - exc);
                if (STORE && act) O::st(fo.ph[3], po.y, p1);
                hz = hz + dmu * p1;
            }
        }
        if ((CPML & 4) && lz.sh >= 0 && (WX || WY)) {
            if constexpr (WX) {
                const T p0 = psi_step(ps.z0, lz.bh, lz.chh, eyk - eyc);
                if (STORE && act) O::st(fo.ph[4], po.z, p0);
                hx = hx + dmu * p0;
            }
            if constexpr (WY) {
                const T p1 = psi_step(ps.z1, lz.bh, lz.chh, exc - exk);
                if (STORE && act) O::st(fo.ph[5], po.z, p1);
                hy = hy + dmu * p1;
            }
        }
    }
    if constexpr (TF) {
        const TfsfArgs& t = fo.t;
        if (i >= t.i0 && i <= t.i1 - 1) {
            if (WY && j >= t.j0 && j <= t.j1 && tl.cy != 0.0) hy = (T)((double)hy + tl.cy);
            if (WZ && (j == t.j0 - 1 || j == t.j1) && tl.e1 != 0.0) {
                const double c = t.dt_mu0 * (double)ldc(a.den_hy, j) * tl.e1;
                hz = (T)((double)hz + (j == t.j1 ? c : -c));
            }
        }
    }
}

// The H update of a halo cell: its psi read here, nothing stored.
template <typename T, int CPML, bool TF, bool WX, bool WY, bool WZ>
__device__ __forceinline__ void h_halo(const YeeArgs<T>& a, const FusedOut<T>& fo, int i, int j, int k, T cx,
                                       const Coef<T>& px, const Coef<T>& ry, const Coef<T>& lz, const TfLane& tl, T ezc,
                                       T ezj, T ezn, T eyc, T eyk, T eyn, T exc, T exk, T exj, T& hx, T& hy, T& hz) {
    const PsiOff po = psi_off<T, false>(a, i, j, k, px, ry, lz);
    const Psi<T> ps = psi_read<T, CPML, false, WX, WY, WZ>(a, po, px, ry, lz);
    h_fused<T, CPML, TF, WX, WY, WZ, false>(a, fo, i, j, k, cx, px, ry, lz, ps, po, tl, ezc, ezj, ezn, eyc, eyk, eyn,
                                            exc, exk, exj, hx, hy, hz);
}

// E^{n+1} of an owned cell (e in: E^n): e_bulk, then the CPML psi of E in
// place (x, y, z; ps: psi of E at n read by psi_read at po); a cell of a
// dispersive class (own false) stores nothing.
template <typename T, int CPML, int MED>
__device__ __forceinline__ void e_fused(const YeeArgs<T>& a, int tg, bool own, T dex, T cex, const Coef<T>& px,
                                        const Coef<T>& ry, const Coef<T>& lz, const Psi<T>& ps, const PsiOff& po,
                                        T hzc, T hzj, T hzi, T hyc, T hyk, T hyi, T hxc, T hxj, T hxk, T& ex, T& ey,
                                        T& ez) {
    using O = Off<T, true>;
    const T dep = a.dt_eps0;
    e_bulk<T, MED>(a, tg, dep, hzc, hzj, hzi, hyc, hyk, hyi, hxc, hxj, hxk, dex, cex, ry.de, lz.de, ry.ce, lz.ce, ex,
                   ey, ez);
    if constexpr (CPML != 0) {
        if ((CPML & 1) && px.se >= 0) {
            const T p0 = psi_step(ps.x0, px.be, px.cee, hzi - hzc);
            const T p1 = psi_step(ps.x1, px.be, px.cee, hyc - hyi);
            if (own) { O::st(a.pe[0], po.x, p0); O::st(a.pe[1], po.x, p1); }
            ey = ey + dep * p0; ez = ez + dep * p1;
        }
        if ((CPML & 2) && ry.se >= 0) {
            const T p0 = psi_step(ps.y0, ry.be, ry.cee, hzc - hzj);
            const T p1 = psi_step(ps.y1, ry.be, ry.cee, hxj - hxc);
            if (own) { O::st(a.pe[2], po.y, p0); O::st(a.pe[3], po.y, p1); }
            ex = ex + dep * p0; ez = ez + dep * p1;
        }
        if ((CPML & 4) && lz.se >= 0) {
            const T p0 = psi_step(ps.z0, lz.be, lz.cee, hyk - hyc);
            const T p1 = psi_step(ps.z1, lz.be, lz.cee, hxc - hxk);
            if (own) { O::st(a.pe[4], po.z, p0); O::st(a.pe[5], po.z, p1); }
            ex = ex + dep * p0; ey = ey + dep * p1;
        }
    }
}

// The fused step over the tiles of fo's boxes.  Workgroups of four
// wavefronts, each its own tile (z tile tz, row block ty, x chunk tx; tz
// fastest).  a: inputs (E^n, H^{n-1/2}, psi of H at n-1/2, psi of E in
// place); fo: outputs.  CPML: the CPML faces the boxes' cells and halo
// cells may lie in (bit 0 x, 1 y, 2 z; the code of the others is left out,
// so a launch over the interior carries no CPML registers).  Requires every
// array below 4 GiB (32-bit byte offsets) and no periodic axis.
template <typename T, int R, int CPML, int MED, bool TF>
__global__ __launch_bounds__(256) void k_step_fused(const YeeArgs<T> a, const FusedOut<T> fo) {
    using O = Off<T, true>;
    const int lane = (int)(threadIdx.x & 63);
    const int wpb = (int)(blockDim.x >> 6);
    int b = (int)blockIdx.x;
    if (fo.tile_order == 1) b = xcd_tile(b, (int)gridDim.x);
    // the wavefront's tile: uniform, so its rows, planes and coefficients sit in scalar registers
    const int t = b * wpb + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (t >= fo.box[fo.nbox - 1].t_end) return;
    int tl = t, bq = 0;
#pragma unroll
    for (int q = 1; q < 4; ++q)
        if (q < fo.nbox && t >= fo.box[q - 1].t_end) { bq = q; tl = t - fo.box[q - 1].t_end; }
    const typename FusedOut<T>::Box& B = fo.box[bq];
    const int tz = B.tz_lo + tl % B.ntz;
    const int r0 = tl / B.ntz;
    const int ty = r0 % B.nty;
    const int tx = r0 / B.nty;
    const int k0 = tz * 64 + 1;
    const int k = k0 + lane;
    const int kc = min(k, a.nz + 1);            // lanes past the grid work on the k = nz+1 cells (kept in
                                                // bounds; their results reach no active lane and are not stored)
    const bool act = k <= a.nz;
    const bool full = k0 + 63 <= a.nz;          // lane 63 is a cell: its k+1 neighbour is loaded
    const int j0 = B.j_lo + ty * R;
    const int nr = min(R, B.j_hi - j0 + 1);
    const int i0 = B.i_lo + tx * B.xc;
    const int i1 = min(i0 + B.xc - 1, B.i_hi);
    const uint32_t S = (uint32_t)sizeof(T);
    const uint32_t sx = (uint32_t)(a.sx * S), sz = (uint32_t)(a.sz * S);
    const T* const cxs = sizeof(T) == 4 ? a.cdhx : a.den_hx;
    const Coef<T> lz = coef_at<T, (CPML & 4) != 0, true, false>(a, a.den_hz, a.cdhz, a.den_ez, a.cdez, kc, a.nz, 16u, 32u);
    const int kh = k0 - 1;                      // lane 0's halo cell (k0 > 1)
    Coef<T> lh = lz;
    if (kh >= 1) lh = coef_at<T, (CPML & 4) != 0, true, true>(a, a.den_hz, a.cdhz, a.den_ez, a.cdez, kh, a.nz, 16u, 32u);
    TfLane tfl{0.0, 0.0}, tfh{0.0, 0.0};        // TF/SF values of the lane's cell and of lane 0's halo cell
    if constexpr (TF) {
        tfl = tf_lane<T, false>(a, fo, k);
        if (kh >= 1) tfh = tf_lane<T, true>(a, fo, kh);
    }
    Coef<T> ry[R + 1];                          // rows j0-1 .. j0+R-1
#pragma unroll
    for (int r = 0; r <= R; ++r)
        ry[r] = coef_at<T, (CPML & 2) != 0, false, true>(a, a.den_hy, a.cdhy, a.den_ey, a.cdey, min(j0 - 1 + r, a.ny + 1), a.ny, 4u, 8u);
    const bool yhalo = j0 > 1;
    uint32_t o = (uint32_t)(((long)i0 * a.sx + (long)j0 * a.sz + kc + a.off) * S);   // (i0, j0, kc)
    // E^n of plane i0, rows j0-1 .. j0+nr-1 (index r+1)
    T exc[R + 1], eyc[R + 1], ezc[R + 1], hyp[R], hzp[R];
#pragma unroll
    for (int r = -1; r < R; ++r) {
        if (r >= nr) break;
        const uint32_t q = o + (uint32_t)(r * (int)sz);
        exc[r + 1] = O::ld(a.ex, q); eyc[r + 1] = O::ld(a.ey, q); ezc[r + 1] = O::ld(a.ez, q);
    }
    // H^{n+1/2} Hy, Hz of plane i0-1 (the x halo): the boundary plane for i0 = 1
    {
        const Coef<T> px = coef_at<T, (CPML & 1) != 0, false, true>(a, a.den_hx, a.cdhx, a.den_ex, a.cdex, i0 - 1, a.nx, 1u,
                                                             2u);
        const T cx = i0 > 1 ? ldc(cxs, i0 - 1) : (T)0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) break;
            const uint32_t p = o + (uint32_t)(r * (int)sz) - sx;
            T hx = (T)0, hy = O::ld(a.hy, p), hz = O::ld(a.hz, p);
            if (i0 > 1) {
                const T exq = O::ld(a.ex, p), eyq = O::ld(a.ey, p), ezq = O::ld(a.ez, p);
                const T exj = O::ld(a.ex, p + sz);
                const T ex63 = (full && lane == 63) ? O::ld(a.ex, p + S) : (T)0;
                const T exk = dpp_shl1(ex63, exq);
                h_halo<T, CPML, TF, false, true, true>(a, fo, i0 - 1, j0 + r, kc, cx, px, ry[r + 1], lz, tfl, ezq,
                                                              (T)0, ezc[r + 1], eyq, (T)0, eyc[r + 1], exq, exk, exj,
                                                              hx, hy, hz);
            }
            hyp[r] = hy; hzp[r] = hz;
        }
    }
    // Each plane in two phases: every load of the plane first (so the
    // wavefront waits on memory once per plane), then the arithmetic.
    for (int i = i0; i <= i1; ++i, o += sx) {
        const T cx = ldc(cxs, i);
        const T dex = ldc(a.den_ex, i), cex = sizeof(T) == 4 ? ldc(a.cdex, i) : (T)0;
        const Coef<T> px = coef_at<T, (CPML & 1) != 0, false, true>(a, a.den_hx, a.cdhx, a.den_ex, a.cdex, i, a.nx, 1u, 2u);
        // -- loads --
        // E^n of plane i+1 (the next plane's current values and this plane's i+1 neighbours)
        T exn[R + 1], eyn[R + 1], ezn[R + 1];
#pragma unroll
        for (int r = -1; r < R; ++r) {
            if (r >= nr) break;
            const uint32_t q = o + (uint32_t)(r * (int)sz) + sx;
            exn[r + 1] = O::ld(a.ex, q); eyn[r + 1] = O::ld(a.ey, q); ezn[r + 1] = O::ld(a.ez, q);
        }
        const uint32_t ob = o + (uint32_t)(nr * (int)sz);          // row j0+nr
        const T ezB = O::ld(a.ez, ob), exB = O::ld(a.ex, ob);
        // halo row j0-1: Hx, Hz (the boundary row for j0 = 1), its psi and Ey(k0+64)
        T hxm = O::ld(a.hx, o - sz), hzm = O::ld(a.hz, o - sz);
        PsiOff pom;
        Psi<T> psm;
        T eym63 = (T)0;
        if (yhalo) {
            pom = psi_off<T, false>(a, i, j0 - 1, kc, px, ry[0], lz);
            psm = psi_read<T, CPML, false, true, false, true>(a, pom, px, ry[0], lz);
            if (full && lane == 63) eym63 = O::ld(a.ey, o - sz + S);
        }
        // per row: H^{n-1/2}, E(k+1) of lane 63, psi of H and E, class tag,
        // and lane 0's inputs for H at k0-1
        T hxo[R], hyo[R], hzo[R], ex63[R], ey63[R];
        T hx0[R], hy0[R], ezh[R], ezjh[R], eznh[R], eyh[R], exh[R];
        int tg[R];
        PsiOff poh[R], poe[R], pol[R];
        Psi<T> psh[R], pse[R], psl[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) break;
            const int j = j0 + r;
            const uint32_t q = o + (uint32_t)(r * (int)sz);
            const Coef<T>& cy = ry[r + 1];
            poh[r] = psi_off<T, false>(a, i, j, kc, px, cy, lz);
            poe[r] = psi_off<T, true>(a, i, j, kc, px, cy, lz);
            psh[r] = psi_read<T, CPML, false, true, true, true>(a, poh[r], px, cy, lz);
            pse[r] = psi_read<T, CPML, true, true, true, true>(a, poe[r], px, cy, lz);
            tg[r] = (MED != 0 && a.tag != nullptr) ? O::tag(a.tag, q) : 0;
            hxo[r] = O::ld(a.hx, q); hyo[r] = O::ld(a.hy, q); hzo[r] = O::ld(a.hz, q);
            ex63[r] = (T)0; ey63[r] = (T)0;
            if (full && lane == 63) { ex63[r] = O::ld(a.ex, q + S); ey63[r] = O::ld(a.ey, q + S); }
            hx0[r] = (T)0; hy0[r] = (T)0;
            ezh[r] = ezjh[r] = eznh[r] = eyh[r] = exh[r] = (T)0;
            if (lane == 0) {
                hx0[r] = O::ld(a.hx, q - S); hy0[r] = O::ld(a.hy, q - S);
                if (kh >= 1) {
                    ezh[r] = O::ld(a.ez, q - S); ezjh[r] = O::ld(a.ez, q - S + sz); eznh[r] = O::ld(a.ez, q - S + sx);
                    eyh[r] = O::ld(a.ey, q - S); exh[r] = O::ld(a.ex, q - S);
                    pol[r] = psi_off<T, false>(a, i, j, kh, px, cy, lh);
                    psl[r] = psi_read<T, CPML, false, true, true, false>(a, pol[r], px, cy, lh);
                }
            }
        }
        // -- arithmetic --
        if (yhalo) {
            const T eyk = dpp_shl1(eym63, eyc[0]);
            T hy = (T)0;
            h_fused<T, CPML, TF, true, false, true, false>(a, fo, i, j0 - 1, kc, cx, px, ry[0], lz, psm, pom, tfl, ezc[0],
                                                          ezc[1], (T)0, eyc[0], eyk, eyn[0], exc[0], (T)0, exc[1],
                                                          hxm, hy, hzm);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) break;
            const int j = j0 + r;
            const uint32_t q = o + (uint32_t)(r * (int)sz);
            const Coef<T>& cy = ry[r + 1];
            T hx = hxo[r], hy = hyo[r], hz = hzo[r];
            const T exk = dpp_shl1(ex63[r], exc[r + 1]), eyk = dpp_shl1(ey63[r], eyc[r + 1]);
            const T ezj = r + 1 < nr ? ezc[r + 1 < R ? r + 2 : R] : ezB;
            const T exj = r + 1 < nr ? exc[r + 1 < R ? r + 2 : R] : exB;
            // lane 0: H of cell k0-1 (the boundary k = 0 for the first z tile)
            T hxk0 = hx0[r], hyk0 = hy0[r];
            if (lane == 0 && kh >= 1) {
                T hzd = (T)0;
                h_fused<T, CPML, TF, true, true, false, false>(a, fo, i, j, kh, cx, px, cy, lh, psl[r], pol[r], tfh, ezh[r],
                                                              ezjh[r], eznh[r], eyh[r], eyc[r + 1], (T)0, exh[r],
                                                              exc[r + 1], (T)0, hxk0, hyk0, hzd);
            }
            // this cell's H^{n+1/2} (psi of H stored by lanes inside the grid)
            h_fused<T, CPML, TF, true, true, true, true>(a, fo, i, j, kc, cx, px, cy, lz, psh[r], poh[r], tfl, ezc[r + 1],
                                                        ezj, ezn[r + 1], eyc[r + 1], eyk, eyn[r + 1], exc[r + 1], exk,
                                                        exj, hx, hy, hz, act);
            const T hxk = dpp_shr1(hxk0, hx), hyk = dpp_shr1(hyk0, hy);
            const bool own = MED != 1 || tg[r] < a.disp_tag0;
            T ex = exc[r + 1], ey = eyc[r + 1], ez = ezc[r + 1];
            e_fused<T, CPML, MED>(a, tg[r], own && act, dex, cex, px, cy, lz, pse[r], poe[r], hz, hzm, hzp[r], hy,
                                  hyk, hyp[r], hx, hxm, hxk, ex, ey, ez);
            if (act) {
                O::st(fo.hx, q, hx); O::st(fo.hy, q, hy); O::st(fo.hz, q, hz);
                if (own) { O::st(fo.ex, q, ex); O::st(fo.ey, q, ey); O::st(fo.ez, q, ez); }
            }
            hyp[r] = hy; hzp[r] = hz;
            hxm = hx; hzm = hz;
        }
#pragma unroll
        for (int r = 0; r <= R; ++r) {
            if (r > nr) break;
            exc[r] = exn[r]; eyc[r] = eyn[r]; ezc[r] = ezn[r];
        }
    }
}

}  // namespace omx
