// omx_fused.hpp -- the fused H+E step (step mode 0) of libomx_fdtd_hip.
//
// One pass per time step: E^n and H^{n-1/2} are read from one buffer set and
// E^{n+1} and H^{n+1/2} written to the other, so each field array is read
// once and written once per step (12 words per cell instead of the 18 of
// separate H and E passes).  The arithmetic of every cell is that of the
// separate kernels (h_tile / e_tile: td_propagate_H_3Dfield and
// td_propagate_E_3Dfield, mxll_3D_mod.f90:505-923), face order included.
//
// A workgroup of NZ x NY wavefronts covers every z cell of the grid (NZ =
// ceil(nz/64) wavefronts of 64 lanes along z) and NY*R rows along y (each
// wavefront R rows), and marches a chunk of xc planes.  Per plane, each
// wavefront loads E^n of the next plane and H^{n-1/2} of its own rows,
// computes H^{n+1/2}, and hands the values its neighbours need to them
// through LDS, one barrier per plane:
//   - the first row's E^n (Ex, Ez) of plane i+1 to the wavefront below (its
//     last row's j+1 neighbour at the next plane),
//   - lane 0's E^n (Ey, Ex) of plane i+1 to the wavefront before in z (its
//     lane 63's k+1 neighbour at the next plane),
//   - the last row's H^{n+1/2} (Hx, Hz) to the wavefront above (its first
//     row's j-1 neighbour),
//   - lane 63's H^{n+1/2} (Hx, Hy) to the wavefront after in z (its lane 0's
//     k-1 neighbour).
// k+-1 inside a wavefront come by DPP lane shifts.  Only the workgroup's
// edges recompute: the row below the first row (H of row j0-1 from inputs;
// j0-1 = 0 is the boundary row, read) and the plane before the chunk (H of
// plane i0-1).  The z edges are the grid's boundary cells (k = 0 read, k =
// nz+1 read), so nothing is recomputed along z.  Halo cells run the same H
// routine as their owners with the CPML psi of H read from the input buffer
// and nothing stored, so their values equal the owner's bit for bit; the psi
// of H is therefore double-buffered like the fields.
//
// Cells of Drude / Drude-Lorentz classes are left to the dispersive pass
// after the kernel (their E and E psi are not stored).  The TF/SF H
// corrections are applied inside (owned and halo cells alike), by the lanes
// and rows on the box's faces only; one extra workgroup advances the
// incident line's H meanwhile (the tiles read only its E).  The E
// corrections are added to the stored E after the kernel, as the separate
// step's k_tfsf_e does.
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

// Workgroup barrier ordering LDS only.  The waves of a fused workgroup
// exchange data through LDS alone (fields, psi of H and the TF/SF line are
// read from one buffer and written to another; the psi of E is read and
// written by its owning lane only), so the barrier need not wait for their
// global loads and stores, as a full __syncthreads would.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Outputs of the fused step (the inputs are the YeeArgs block) and its
// TF/SF data.
template <typename T>
struct FusedOut {
    T *ex, *ey, *ez, *hx, *hy, *hz;   // E^{n+1}, H^{n+1/2}
    T* ph[6];                         // psi of H at n+1/2 (x, y, z slab arrays as YeeArgs::ph)
    const double* ex1;                // TF/SF: the incident line's E^n (null without TF/SF)
    double* hy1;                      // TF/SF: the line's H, advanced to n+1/2 by the last workgroup
    TfsfArgs t;
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

// The CPML psi values of a cell read ahead of their use (zero where the cell
// has no such face): x face (Hy,Hz | Ey,Ez), y face (Hx,Hz | Ex,Ez), z face
// (Hx,Hy | Ex,Ey).  Reading a row's psi together with its fields keeps them
// in one memory round trip instead of one per face.
template <typename T>
struct Psi {
    T x0, x1, y0, y1, z0, z1;
};

// Workgroup tiles of k_step_wg, in boxes numbered one after the other:
// planes i_lo..i_hi and rows j_lo..j_hi in row blocks of NY*R rows (nyb of
// them, fastest) and chunks of xc planes; b_end is one past the box's last
// workgroup.  The host lists the boxes holding CPML faces first, so the
// heaviest tiles start first.
struct WgBox {
    int i_lo, i_hi, j_lo, j_hi, nyb, xc, b_end;
};
constexpr int kWgBoxes = 5;
struct WgPlan {
    WgBox box[kWgBoxes];
    int nbox;
    int nzw, nyw;        // wavefronts per workgroup along z and y
};

// LDS of one workgroup: with TF/SF, each wavefront's per-lane TF/SF value
// (NW x 64 doubles), then two buffers (plane parity) of the row
// hand-overs (NW x 2 components x 64 lanes, H then E) and of the lane
// hand-overs (NW x R rows x 2 components, H then E).
__host__ __device__ inline int wg_lds_elems(int nw, int r) { return 8 * nw * 64 + 8 * nw * r; }
__host__ __device__ inline size_t wg_lds_bytes(int nw, int r, size_t esz, bool tf) {
    return (tf ? (size_t)nw * 64 * sizeof(double) : 0) + esz * (size_t)wg_lds_elems(nw, r);
}

// Element of array p at the uniform byte offset uo plus the lane's byte
// offset v: one 32-bit vector offset from the array's scalar base (adding
// the two parts in 32 bits keeps the compiler from holding 64-bit addresses
// of every array in vector registers across the plane loop).
template <typename T>
__device__ __forceinline__ T ldu(const T* p, uint32_t uo, uint32_t v) {
    return ldb(p, uo + v);
}
template <typename T>
__device__ __forceinline__ void stu(T* p, uint32_t uo, uint32_t v, T x) {
    stb(p, uo + v, x);
}

// Uniform byte offsets of the psi of an x face slot / a y face slot (ny2
// slots per plane) / the z faces at (i, j).
template <typename T>
__device__ __forceinline__ uint32_t uo_psix(const YeeArgs<T>& a, int slot, int j) {
    return ((uint32_t)slot * (uint32_t)a.sx + (uint32_t)j * (uint32_t)a.sz) * (uint32_t)sizeof(T);
}
template <typename T>
__device__ __forceinline__ uint32_t uo_psiy(const YeeArgs<T>& a, int i, int slot, int ny2) {
    return (((uint32_t)i * (uint32_t)ny2 + (uint32_t)slot) * (uint32_t)a.sz) * (uint32_t)sizeof(T);
}
template <typename T>
__device__ __forceinline__ uint32_t uo_psiz(const YeeArgs<T>& a, int i, int j) {
    return (((uint32_t)j * (uint32_t)(a.nx + 2) + (uint32_t)i) * (uint32_t)(a.npml - 1)) * (uint32_t)sizeof(T);
}

// The psi of H (E false: the components WX/WY/WZ need) or of E (E true:
// all) of the faces cell (i,j,k) lies in, zero elsewhere.  The x and y faces
// are uniform over the wavefront (plane, row), the z faces per lane.  kS:
// the lane's byte offset in a row; zS: its z-face psi byte offset.
template <typename T, int CPML, bool E, bool WX, bool WY, bool WZ>
__device__ __forceinline__ Psi<T> psi_wg(const YeeArgs<T>& a, int i, int j, uint32_t kS, uint32_t zS, const Coef<T>& px,
                                         const Coef<T>& ry, const Coef<T>& lz, bool zface) {
    Psi<T> v;
    v.x0 = v.x1 = v.y0 = v.y1 = v.z0 = v.z1 = (T)0;
    if constexpr (CPML != 0) {
        T* const* ps = E ? a.pe : a.ph;
        const int sx_ = E ? px.se : px.sh, sy_ = E ? ry.se : ry.sh, sz_ = E ? lz.se : lz.sh;
        if ((CPML & 1) && sx_ >= 0) {
            const uint32_t uo = uo_psix(a, sx_, j);
            if (E || WY) v.x0 = ldu(ps[0], uo, kS);
            if (E || WZ) v.x1 = ldu(ps[1], uo, kS);
        }
        if ((CPML & 2) && sy_ >= 0) {
            const uint32_t uo = uo_psiy(a, i, sy_, E ? 2 * a.npml : 2 * (a.npml - 1));
            if (E || WX) v.y0 = ldu(ps[2], uo, kS);
            if (E || WZ) v.y1 = ldu(ps[3], uo, kS);
        }
        if ((CPML & 4) && zface && sz_ >= 0) {
            const uint32_t uo = uo_psiz(a, i, j);
            if (E || WX) v.z0 = ldu(ps[4], uo, zS);
            if (E || WY) v.z1 = ldu(ps[5], uo, zS);
        }
    }
    return v;
}

// H^{n+1/2} of the components WX/WY/WZ of cell (i,j,k) with the arithmetic
// of h_tile: the bulk curl, the CPML psi of the x, y and z faces in that
// order (ps: psi of H at n-1/2 from psi_wg), then the TF/SF correction where
// the plane, row and lane lie on a face of the box (tfv: the lane's value of
// this step in LDS).  STORE writes the new psi to fo.ph where act is set.
template <typename T, int CPML, bool TF, bool WX, bool WY, bool WZ, bool STORE>
__device__ __forceinline__ void h_wg(const YeeArgs<T>& a, const FusedOut<T>& fo, const double* tfv, int i, int j, int k,
                                     uint32_t kS, uint32_t zhS, T cx, const Coef<T>& px, const Coef<T>& ry,
                                     const Coef<T>& lz, const Psi<T>& ps, T ezc, T ezj, T ezn, T eyc, T eyk, T eyn, T exc,
                                     T exk, T exj, T& hx, T& hy, T& hz, bool act) {
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
        if ((CPML & 1) && (WY || WZ) && px.sh >= 0) {
            const uint32_t uo = uo_psix(a, px.sh, j);
            if constexpr (WY) {
                const T p0 = psi_step(ps.x0, px.bh, px.chh, ezn - ezc);
                if (STORE && act) stu(fo.ph[0], uo, kS, p0);
                hy = hy + dmu * p0;
            }
            if constexpr (WZ) {
                const T p1 = psi_step(ps.x1, px.bh, px.chh, eyc - eyn);
                if (STORE && act) stu(fo.ph[1], uo, kS, p1);
                hz = hz + dmu * p1;
            }
        }
        if ((CPML & 2) && (WX || WZ) && ry.sh >= 0) {
            const uint32_t uo = uo_psiy(a, i, ry.sh, 2 * (a.npml - 1));
            if constexpr (WX) {
                const T p0 = psi_step(ps.y0, ry.bh, ry.chh, ezc - ezj);
                if (STORE && act) stu(fo.ph[2], uo, kS, p0);
                hx = hx + dmu * p0;
            }
            if constexpr (WZ) {
                const T p1 = psi_step(ps.y1, ry.bh, ry.chh, exj - exc);
                if (STORE && act) stu(fo.ph[3], uo, kS, p1);
                hz = hz + dmu * p1;
            }
        }
        if ((CPML & 4) && (WX || WY) && lz.sh >= 0) {
            const uint32_t uo = uo_psiz(a, i, j);
            if constexpr (WX) {
                const T p0 = psi_step(ps.z0, lz.bh, lz.chh, eyk - eyc);
                if (STORE && act) stu(fo.ph[4], uo, zhS, p0);
                hx = hx + dmu * p0;
            }
            if constexpr (WY) {
                const T p1 = psi_step(ps.z1, lz.bh, lz.chh, exc - exk);
                if (STORE && act) stu(fo.ph[5], uo, zhS, p1);
                hy = hy + dmu * p1;
            }
        }
    }
    if constexpr (TF) {
        const TfsfArgs& t = fo.t;
        if (i >= t.i0 && i <= t.i1 - 1) {
            if (WY && j >= t.j0 && j <= t.j1 && (k == t.k0 - 1 || k == t.k1)) hy = (T)((double)hy + tfv[0]);
            if (WZ && (j == t.j0 - 1 || j == t.j1) && k >= t.k0 && k <= t.k1) {
                const double e1 = fo.ex1[k - t.kb];
                if (e1 != 0.0) {
                    const double c = t.dt_mu0 * (double)ldc(a.den_hy, j) * e1;
                    hz = (T)((double)hz + (j == t.j1 ? c : -c));
                }
            }
        }
    }
}

// E^{n+1} of an owned cell with the arithmetic of e_tile: e_bulk, then the
// CPML psi of E in place (x, y, z faces; ps: psi of E at n from psi_wg);
// own false (a cell of a dispersive class, or a lane past the grid) stores
// no psi.
template <typename T, int CPML, int MED>
__device__ __forceinline__ void e_wg(const YeeArgs<T>& a, int i, int j, uint32_t kS, uint32_t zeS, int tg, bool own,
                                     T dex, T cex, const Coef<T>& px, const Coef<T>& ry, const Coef<T>& lz, const Psi<T>& ps, T hzc, T hzj, T hzi, T hyc, T hyk, T hyi,
                                     T hxc, T hxj, T hxk, T& ex, T& ey, T& ez) {
    const T dep = a.dt_eps0;
    // a cell of a dispersive class is updated by the dispersive pass (nothing
    // stored here): the vacuum form spares its class lookups
    const int tgb = MED == 1 && tg >= a.disp_tag0 ? 0 : tg;
    e_bulk<T, MED>(a, tgb, dep, hzc, hzj, hzi, hyc, hyk, hyi, hxc, hxj, hxk, dex, cex, ry.de, lz.de, ry.ce, lz.ce, ex,
                   ey, ez);
    if constexpr (CPML != 0) {
        if ((CPML & 1) && px.se >= 0) {
            const uint32_t uo = uo_psix(a, px.se, j);
            const T p0 = psi_step(ps.x0, px.be, px.cee, hzi - hzc);
            const T p1 = psi_step(ps.x1, px.be, px.cee, hyc - hyi);
            if (own) { stu(a.pe[0], uo, kS, p0); stu(a.pe[1], uo, kS, p1); }
            ey = ey + dep * p0; ez = ez + dep * p1;
        }
        if ((CPML & 2) && ry.se >= 0) {
            const uint32_t uo = uo_psiy(a, i, ry.se, 2 * a.npml);
            const T p0 = psi_step(ps.y0, ry.be, ry.cee, hzc - hzj);
            const T p1 = psi_step(ps.y1, ry.be, ry.cee, hxj - hxc);
            if (own) { stu(a.pe[2], uo, kS, p0); stu(a.pe[3], uo, kS, p1); }
            ex = ex + dep * p0; ez = ez + dep * p1;
        }
        if ((CPML & 4) && lz.se >= 0) {
            const uint32_t uo = uo_psiz(a, i, j);
            const T p0 = psi_step(ps.z0, lz.be, lz.cee, hyk - hyc);
            const T p1 = psi_step(ps.z1, lz.be, lz.cee, hxc - hxk);
            if (own) { stu(a.pe[4], uo, zeS, p0); stu(a.pe[5], uo, zeS, p1); }
            ex = ex + dep * p0; ey = ey + dep * p1;
        }
    }
}

// FP64 workgroups hold at most 8 wavefronts (the host's shape rule), so the
// FP64 kernels are bounded at 512 threads: a lane may then take up to 256
// VGPRs, and the CPML variants (137-218) do not spill as they did at 128.
template <typename T, int R, int CPML, int MED, bool TF>
__global__ __launch_bounds__(sizeof(T) == 8 ? 512 : 1024) void k_step_wg(const YeeArgs<T> a, const FusedOut<T> fo, const WgPlan P) {
    extern __shared__ __attribute__((aligned(16))) char wg_lds_raw[];
    const int NZ = P.nzw, NY = P.nyw, NW = NZ * NY;
    const int lane = (int)(threadIdx.x & 63);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int zw = wv % NZ, yw = wv / NZ;
    double* const tfv = reinterpret_cast<double*>(wg_lds_raw) + wv * 64 + lane;   // TF: this lane's value
    T* const lds = reinterpret_cast<T*>(wg_lds_raw + (TF ? NW * 64 * sizeof(double) : 0));
    const int b = (int)blockIdx.x;
    if constexpr (TF) {
        // the workgroup after the last box: the incident line's H half step
        // (n-1/2 -> n+1/2), which the E corrections after the kernel read
        int nb = P.box[0].b_end;
#pragma unroll
        for (int q = 1; q < kWgBoxes; ++q)
            if (q < P.nbox) nb = P.box[q].b_end;
        if (b == nb) {
            const TfsfArgs& t = fo.t;
            for (int m = (int)threadIdx.x; m < t.n1 - 1; m += (int)blockDim.x)
                fo.hy1[m] = fo.hy1[m] + t.dt_mu0 * t.inv_dr * (fo.ex1[m] - fo.ex1[m + 1]);
            return;
        }
    }
    int bq = 0, bl = b;
#pragma unroll
    for (int q = 1; q < kWgBoxes; ++q)
        if (q < P.nbox && b >= P.box[q - 1].b_end) { bq = q; bl = b - P.box[q - 1].b_end; }
    const WgBox& B = P.box[bq];
    const int yb = bl % B.nyb, xb = bl / B.nyb;
    const int j0 = B.j_lo + (yb * NY + yw) * R;
    const int nr = max(0, min(R, B.j_hi - j0 + 1));        // rows of this wavefront (0: it only joins the barriers)
    const int i0 = B.i_lo + xb * B.xc, i1 = min(i0 + B.xc - 1, B.i_hi);
    const int k0 = 1 + 64 * zw, k = k0 + lane;
    const int kc = min(k, a.nz + 1);                        // lanes past the grid read the k = nz+1 cells
    const bool act = k <= a.nz;
    const bool full = k0 + 63 <= a.nz;                      // lane 63 is a cell: its k+1 neighbour is read
    const bool zlo = zw == 0, zhi = zw == NZ - 1;
    // the wavefront's k range meets a z face (the high face may straddle two
    // wavefronts when nz is not a multiple of 64)
    const bool zface = (CPML & 4) && (k0 <= a.npml || k0 + 63 >= a.nz + 1 - a.npml);
    const bool top_lds = yw < NY - 1 && j0 + R <= B.j_hi;   // the wavefront above has rows: j+1 of the last row from it
    const bool ybot = yw == 0;
    const bool yhalo = ybot && j0 > 1;                      // row j0-1 recomputed (j0 = 1: the boundary row, read)
    T* const hrow = lds;                                    // [2][NW][2][64]: Hx, Hz of the last row
    T* const erow = lds + 4 * NW * 64;                      // [2][NW][2][64]: Ex, Ez of the first row (next plane)
    T* const hlz = lds + 8 * NW * 64;                       // [2][NW][R][2]: lane 63's Hx, Hy
    T* const elz = hlz + 4 * NW * R;                        // [2][NW][R][2]: lane 0's Ey, Ex (next plane)
    constexpr uint32_t S = (uint32_t)sizeof(T);
    const uint32_t kS = (uint32_t)(kc + a.off) * S;         // the lane's byte offset in a row
    const uint32_t rS = (uint32_t)a.sz * S, pS = (uint32_t)a.sx * S;
    const T* const cxs = sizeof(T) == 4 ? a.cdhx : a.den_hx;
    const Coef<T> lz = coef_at<T, (CPML & 4) != 0, true, false>(a, a.den_hz, a.cdhz, a.den_ez, a.cdez, kc, a.nz, 16u, 32u);
    const uint32_t zhS = (uint32_t)max(lz.sh, 0) * S, zeS = (uint32_t)max(lz.se, 0) * S;
    if constexpr (TF) {
        // the lane's TF/SF correction of this step: Hy on the box's z faces
        // (k = k0-1: + dt/mu0 den_hz(k) Ex_inc(k0), k = k1: - ... Ex_inc(k1));
        // read only by the lanes on those faces
        const TfsfArgs& t = fo.t;
        double cy = 0.0;
        if (k == t.k0 - 1 || k == t.k1) {
            const bool hi = k == t.k1;
            const double c = t.dt_mu0 * (double)gld(a.den_hz, k) * fo.ex1[(hi ? t.k1 : t.k0) - t.kb];
            cy = hi ? -c : c;
        }
        tfv[0] = cy;
    }
    Coef<T> ry[R + 1];                                      // rows j0-1 .. j0+R-1
    T exc[R], eyc[R], ezc[R], hyp[R], hzp[R];
    T exh = (T)0, eyh = (T)0, ezh = (T)0;                   // E^n of row j0-1 (yhalo)
    if (nr > 0) {
#pragma unroll
        for (int r = 0; r <= R; ++r)
            ry[r] = coef_at<T, (CPML & 2) != 0, false, true>(a, a.den_hy, a.cdhy, a.den_ey, a.cdey,
                                                            min(max(j0 - 1 + r, 0), a.ny + 1), a.ny, 4u, 8u);
        const uint32_t u0 = (uint32_t)i0 * pS + (uint32_t)j0 * rS;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) break;
            const uint32_t u = u0 + (uint32_t)r * rS;
            exc[r] = ldu(a.ex, u, kS); eyc[r] = ldu(a.ey, u, kS); ezc[r] = ldu(a.ez, u, kS);
        }
        if (yhalo) { exh = ldu(a.ex, u0 - rS, kS); eyh = ldu(a.ey, u0 - rS, kS); ezh = ldu(a.ez, u0 - rS, kS); }
        // H^{n+1/2} Hy, Hz of plane i0-1 (i0 = 1: the boundary plane, read)
        const Coef<T> px = coef_at<T, (CPML & 1) != 0, false, true>(a, a.den_hx, a.cdhx, a.den_ex, a.cdex, i0 - 1, a.nx,
                                                                   1u, 2u);
        const T cx = i0 > 1 ? ldc(cxs, i0 - 1) : (T)0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r >= nr) break;
            const uint32_t u = u0 + (uint32_t)r * rS - pS;
            T hx = (T)0, hy = ldu(a.hy, u, kS), hz = ldu(a.hz, u, kS);
            if (i0 > 1) {
                const T exq = ldu(a.ex, u, kS), eyq = ldu(a.ey, u, kS), ezq = ldu(a.ez, u, kS);
                const T exj = ldu(a.ex, u + rS, kS);
                const T ex63 = (full && lane == 63) ? ldu(a.ex, u, kS + S) : (T)0;
                const T exk = dpp_shl1(ex63, exq);
                const Psi<T> ps = psi_wg<T, CPML, false, false, true, true>(a, i0 - 1, j0 + r, kS, zhS, px, ry[r + 1], lz,
                                                                          zface);
                h_wg<T, CPML, TF, false, true, true, false>(a, fo, tfv, i0 - 1, j0 + r, k, kS, zhS, cx, px, ry[r + 1], lz, ps,
                                                           ezq, (T)0, ezc[r], eyq, (T)0, eyc[r], exq, exk, exj, hx, hy,
                                                           hz, false);
            }
            hyp[r] = hy; hzp[r] = hz;
        }
        const int pb = i0 & 1;
        erow[((pb * NW + wv) * 2 + 0) * 64 + lane] = exc[0];
        erow[((pb * NW + wv) * 2 + 1) * 64 + lane] = ezc[0];
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r >= nr) break;
                elz[((pb * NW + wv) * R + r) * 2 + 0] = eyc[r];
                elz[((pb * NW + wv) * R + r) * 2 + 1] = exc[r];
            }
        }
    }
    lds_barrier();
    for (int i = i0; i <= i1; ++i) {
        const int pb = i & 1, nb = pb ^ 1;
        const uint32_t u0 = (uint32_t)i * pS + (uint32_t)max(j0, 0) * rS;
        T exn[R], eyn[R], ezn[R], hx[R], hy[R], hz[R], ex63[R], ey63[R], hx0[R], hy0[R];
        Psi<T> psh[R], pse[R];
        int tg[R];
        T exnh = (T)0, eynh = (T)0, eznh = (T)0, hxm = (T)0, hzm = (T)0, eym63 = (T)0, ezB = (T)0, exB = (T)0;
        T dex = (T)0, cex = (T)0;
        Coef<T> px;
        if (nr > 0) {
            const T cx = ldc(cxs, i);
            dex = ldc(a.den_ex, i);
            cex = sizeof(T) == 4 ? ldc(a.cdex, i) : (T)0;
            px = coef_at<T, (CPML & 1) != 0, false, true>(a, a.den_hx, a.cdhx, a.den_ex, a.cdex, i, a.nx, 1u, 2u);
            // -- loads: E^n of plane i+1, H^{n-1/2}, psi and tags of the rows; the edges' inputs --
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r >= nr) break;
                const uint32_t u = u0 + (uint32_t)r * rS;
                exn[r] = ldu(a.ex, u + pS, kS); eyn[r] = ldu(a.ey, u + pS, kS); ezn[r] = ldu(a.ez, u + pS, kS);
                hx[r] = ldu(a.hx, u, kS); hy[r] = ldu(a.hy, u, kS); hz[r] = ldu(a.hz, u, kS);
                psh[r] = psi_wg<T, CPML, false, true, true, true>(a, i, j0 + r, kS, zhS, px, ry[r + 1], lz, zface);
                pse[r] = psi_wg<T, CPML, true, true, true, true>(a, i, j0 + r, kS, zeS, px, ry[r + 1], lz, zface);
                tg[r] = (MED != 0 && a.tag != nullptr) ? (int)ldb(a.tag, u / S + (uint32_t)(kc + a.off)) : 0;
                ex63[r] = ey63[r] = hx0[r] = hy0[r] = (T)0;
                if (zhi && full && lane == 63) { ex63[r] = ldu(a.ex, u, kS + S); ey63[r] = ldu(a.ey, u, kS + S); }
                if (zlo && lane == 0) { hx0[r] = ldu(a.hx, u, kS - S); hy0[r] = ldu(a.hy, u, kS - S); }
            }
            Psi<T> psm;
            if (yhalo) {
                exnh = ldu(a.ex, u0 - rS + pS, kS); eynh = ldu(a.ey, u0 - rS + pS, kS); eznh = ldu(a.ez, u0 - rS + pS, kS);
                if (full && lane == 63) eym63 = ldu(a.ey, u0 - rS, kS + S);
                psm = psi_wg<T, CPML, false, true, false, true>(a, i, j0 - 1, kS, zhS, px, ry[0], lz, zface);
            }
            if (ybot) { hxm = ldu(a.hx, u0 - rS, kS); hzm = ldu(a.hz, u0 - rS, kS); }
            if (!top_lds) {
                ezB = ldu(a.ez, u0 + (uint32_t)nr * rS, kS);
                exB = ldu(a.ex, u0 + (uint32_t)nr * rS, kS);
            }
            // -- hand E^n of plane i+1 to the wavefronts below and before; take this plane's from above and after --
            erow[((nb * NW + wv) * 2 + 0) * 64 + lane] = exn[0];
            erow[((nb * NW + wv) * 2 + 1) * 64 + lane] = ezn[0];
            if (lane == 0) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r >= nr) break;
                    elz[((nb * NW + wv) * R + r) * 2 + 0] = eyn[r];
                    elz[((nb * NW + wv) * R + r) * 2 + 1] = exn[r];
                }
            }
            if (top_lds) {
                exB = erow[((pb * NW + wv + NZ) * 2 + 0) * 64 + lane];
                ezB = erow[((pb * NW + wv + NZ) * 2 + 1) * 64 + lane];
            }
            if (!zhi) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r >= nr) break;
                    ey63[r] = elz[((pb * NW + wv + 1) * R + r) * 2 + 0];
                    ex63[r] = elz[((pb * NW + wv + 1) * R + r) * 2 + 1];
                }
            }
            // -- H^{n+1/2} of the rows (and of row j0-1) --
            if (yhalo) {
                const T eyk = dpp_shl1(eym63, eyh);
                T hyd = (T)0;
                h_wg<T, CPML, TF, true, false, true, false>(a, fo, tfv, i, j0 - 1, k, kS, zhS, cx, px, ry[0], lz, psm, ezh,
                                                           ezc[0], (T)0, eyh, eyk, eynh, exh, (T)0, exc[0], hxm, hyd, hzm,
                                                           false);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r >= nr) break;
                const T ezj = r + 1 < nr ? ezc[r + 1 < R ? r + 1 : 0] : ezB;
                const T exj = r + 1 < nr ? exc[r + 1 < R ? r + 1 : 0] : exB;
                const T exk = dpp_shl1(ex63[r], exc[r]), eyk = dpp_shl1(ey63[r], eyc[r]);
                h_wg<T, CPML, TF, true, true, true, true>(a, fo, tfv, i, j0 + r, k, kS, zhS, cx, px, ry[r + 1], lz, psh[r],
                                                         ezc[r], ezj, ezn[r], eyc[r], eyk, eyn[r], exc[r], exk, exj,
                                                         hx[r], hy[r], hz[r], act);
            }
            // -- hand H^{n+1/2} to the wavefronts above and after --
            if (nr == R) {
                hrow[((pb * NW + wv) * 2 + 0) * 64 + lane] = hx[R - 1];
                hrow[((pb * NW + wv) * 2 + 1) * 64 + lane] = hz[R - 1];
            }
            if (lane == 63) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r >= nr) break;
                    hlz[((pb * NW + wv) * R + r) * 2 + 0] = hx[r];
                    hlz[((pb * NW + wv) * R + r) * 2 + 1] = hy[r];
                }
            }
        }
        lds_barrier();
        if (nr > 0) {
            if (!ybot) {
                hxm = hrow[((pb * NW + wv - NZ) * 2 + 0) * 64 + lane];
                hzm = hrow[((pb * NW + wv - NZ) * 2 + 1) * 64 + lane];
            }
            if (!zlo) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (r >= nr) break;
                    hx0[r] = hlz[((pb * NW + wv - 1) * R + r) * 2 + 0];
                    hy0[r] = hlz[((pb * NW + wv - 1) * R + r) * 2 + 1];
                }
            }
            // -- E^{n+1} of the rows, stores --
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r >= nr) break;
                const uint32_t u = u0 + (uint32_t)r * rS;
                const T hxk = dpp_shr1(hx0[r], hx[r]), hyk = dpp_shr1(hy0[r], hy[r]);
                const T hzj = r > 0 ? hz[r > 0 ? r - 1 : 0] : hzm;
                const T hxj = r > 0 ? hx[r > 0 ? r - 1 : 0] : hxm;
                const bool own = MED != 1 || tg[r] < a.disp_tag0;
                T ex = exc[r], ey = eyc[r], ez = ezc[r];
                e_wg<T, CPML, MED>(a, i, j0 + r, kS, zeS, tg[r], own && act, dex, cex, px, ry[r + 1], lz, pse[r], hz[r], hzj, hzp[r], hy[r], hyk, hyp[r], hx[r], hxj, hxk, ex, ey, ez);
                if (act) {
                    stu(fo.hx, u, kS, hx[r]); stu(fo.hy, u, kS, hy[r]); stu(fo.hz, u, kS, hz[r]);
                    if (own) { stu(fo.ex, u, kS, ex); stu(fo.ey, u, kS, ey); stu(fo.ez, u, kS, ez); }
                }
                hyp[r] = hy[r]; hzp[r] = hz[r];
                exc[r] = exn[r]; eyc[r] = eyn[r]; ezc[r] = ezn[r];
            }
            if (yhalo) { exh = exnh; eyh = eynh; ezh = eznh; }
        }
    }
}


// After the workgroup-tile fused step: the cells of dispersive classes
// (workgroups [0, nd), one cell per thread, as k_e_dispersive; STEPPED when
// the fused step ran without its media variant and so already advanced
// their E psi) and, when no such cell lies on the box's E faces, the TF/SF
// E step (the remaining workgroups: k_tfsf_e's corrections to E^{n+1} and
// the 1D line's E half step with the source).  The two parts touch disjoint
// data.  LF: each cell loads its three components' state and its E psi
// before any update (for lists short enough that the pass is bound by
// memory latency, not bandwidth: the extra registers cost waves).
template <typename T, bool CPML, bool B32, bool STEPPED, bool LF>
__global__ __launch_bounds__(256) void k_wg_post(const YeeArgs<T> a, T* __restrict__ exo, T* __restrict__ eyo,
                                                 T* __restrict__ ezo, const long* __restrict__ list, int n, int nd,
                                                 TfsfArgs t, double* __restrict__ ex1, const double* __restrict__ hy1,
                                                 double src) {
    const int b = (int)blockIdx.x;
    if (b < nd) {
        const int p = b * (int)blockDim.x + (int)threadIdx.x;
        if (p < n) e_dispersive_cell<T, CPML, B32, STEPPED, LF>(a, exo, eyo, ezo, list[p], p);
        return;
    }
    tfsf_e_item<T>(a, exo, ezo, t, ex1, hy1, src, (b - nd) * (int)blockDim.x + (int)threadIdx.x);
}

}  // namespace omx
