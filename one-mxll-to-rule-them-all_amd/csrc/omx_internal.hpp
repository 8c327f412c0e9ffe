// omx_internal.hpp -- private types of libomx_fdtd_hip: device layout, kernel
// argument blocks and the context behind the C-ABI handle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "omx_fdtd.h"

namespace omx {

// Physical constants in atomic units, with the values of the reference's
// constants module (src/FDTD/constants_mod.f90:15-47).  Several reference
// literals are default-REAL (single precision) and are kept at their
// single-rounded values here so that derived quantities agree bit for bit.
constexpr double kPi        = 3.141592653589793;
constexpr double kC0        = 137.0;
constexpr double kEps0      = 1.0 / (4.0 * kPi);
constexpr double kMu0       = 1.0 / (kEps0 * (kC0 * kC0));
constexpr double kAlphaCPML = (double)0.05f;   // constants_mod.f90:27 (REAL(4) literal)
constexpr double kKappaCPML = 5.0;
constexpr int    kGradeM    = 3;               // polynomial grading order m
constexpr int    kGradeMa   = 1;               // alpha grading order ma
constexpr double kSecToAu   = 1.0 / 2.4188843265864e-17;
constexpr double kCToAu     = 1.0 / 1.602176634e-19;
constexpr double kKgToAu    = 1.0 / 9.1093837139e-31;
constexpr double kMToAu     = 1.0 / 5.29177210544e-11;
constexpr double kEvToAu    = (double)(1.0f / 27.2114f);  // constants_mod.f90:37, REAL(4) quotient

// Device layout of every field-shaped array: index (i,j,k), each in 0..n+1,
// z fastest; rows padded so that k = 1 starts on a 128-byte boundary.
struct Layout {
    int nx, ny, nz;
    int sy;        // ny + 2
    int sz;        // padded row length in elements
    long sx;       // sy * sz
    int off;       // element offset of (0,0,0)
    long nalloc;   // elements per array
    __host__ __device__ inline long at(int i, int j, int k) const {
        return (long)i * sx + (long)j * sz + k + off;
    }
};

template <typename T>
struct MedDev {
    int kind, np;
    T cdt;                     // dt_eps0 / eps_r for dielectric cells
    T C1, C2, C3, C4, A1, A2;  // FP32 keeps C1-1 and A1-1 (delta form)
    T al[OMX_MAX_POLES], bt[OMX_MAX_POLES], ga[OMX_MAX_POLES];  // FP32: alpha-2, beta+1
    T B1[OMX_MAX_POLES], B2[OMX_MAX_POLES];
};

template <typename T>
struct YeeArgs {
    T *ex, *ey, *ez, *hx, *hy, *hz;
    const T *den_ex, *den_ey, *den_ez, *den_hx, *den_hy, *den_hz;  // 1-based
    const T *bh, *chd, *be, *ced;                                  // 1-based; FP32: b-1
    // FP32 only: dt_eps0*den_e* and dt_mu0*den_h* folded into one coefficient (1-based)
    const T *cdex, *cdey, *cdez, *cdhx, *cdhy, *cdhz;
    // H psi: psix_Hyx psix_Hzx psiy_Hxy psiy_Hzy psiz_Hxz psiz_Hyz
    T *ph[6];
    // E psi: psix_Eyx psix_Ezx psiy_Exy psiy_Ezy psiz_Exz psiz_Eyz
    T *pe[6];
    const uint8_t* tag;
    const int8_t* cls;
    const MedDev<T>* med;
    // Drude / DL state of the listed dispersive cells, one block: component c
    // at dstate + c * dcomp holds E_old, PD, then maxp PL planes and maxp
    // PL_old planes, each pstride (the list's length) elements
    T* dstate;
    long dcomp;
    long pstride;
    int maxp;
    long disp0;                // position of the grid's first dispersive cell in the list the kernel reads
    T dt_mu0, dt_eps0;
    int nx, ny, nz, sy, sz;
    long sx;
    int off;
    int npml;
    unsigned cpml;             // bit f set: CPML on face f (x-lo, x-hi, y-lo, y-hi, z-lo, z-hi)
    int disp_tag0;             // tags >= disp_tag0: classes with a Drude / Drude-Lorentz component
    int ntz, nty, xchunk, rows;
    int tile_order;            // 0: tiles in block order, 1: one contiguous slab of tiles per XCD,
                               // 2: whole row blocks (all x chunks) dealt to the XCDs
};

struct SourceDev {
    int dir;
    double j_amp, w0, t0, tau, t_init, t_final, phase;
    int npts;
    long* d_off;        // field offsets of the kernel points
    double* d_ka;       // ker_mat at the point
    double* d_kb;       // ker_mat one step along dir (0 past the kernel edge)
    std::vector<int> h_ijk;          // the points' cells (i, j, k), 3 per point
    std::vector<double*> h_sh;       // per point: its near-field shadow element, or null
    double** d_sh;                   // device copy of h_sh (null when no point has one)
};

// One cell that quantum systems write J into.  sys / coef: the last writer
// of the step (J = coef * dPt/dt of system sys, coef = density * kernel
// weight); mode 0: the last q-group writes the cell, and J_old is what the
// groups before it left there (prev / coef_prev: their last writer, -1: the
// cell's J of the previous quantum step); mode 1: only earlier groups write
// it, so the last group's J_old <- J copy leaves J_old = J and dJ = 0.
struct QSlot {
    int sys, prev, mode, pad;
    double coef, coef_prev;
};

// A near-field patch box of an FP32 grid (kernels: k_patch).
struct PatchBox {
    int lo[3], w[3];            // first cell (1-based) and extent per axis
    long first;                 // first cell of the box in the launch's flat index
    double* pe;                 // shadow of E: component c at pe + c * cs, cell ((i*w1)+j)*w2+k
    double* ph;                 // shadow of H, same layout
    long cs;
    const double* den[6];       // FP64 den_ex den_ey den_ez den_hx den_hy den_hz (1-based)
    const double* cdt;          // per class and component: dt_eps0 / eps_r (dt_eps0 in vacuum)
    double dt_mu0, dt_eps0;
    int g;                      // argument block of the box's grid (batched launches)
    // cells of Drude / Drude-Lorentz classes in the box: their position in the
    // grid's dispersive list (-1: other cells), per box cell; null if none
    const int* dl;
    // FP64 Drude / DL state of the box cells: variable v (E_old, PD, maxp PL,
    // maxp PL_old) of component c at ds + (c * nv + v) * cs, cell as for pe
    double* ds;
    int nv;                     // 2 + 2 * maxp
    const omx_medium_coeffs* mco;   // FP64 update coefficients per medium
};

struct DetectorHost {
    int type, field, nd;
    long first;         // first sample index in the flattened record row
};

}  // namespace omx

struct omx3d_ctx {
    omx3d_desc d;
    int device = 0;     // the GPU current at omx3d_create: every buffer of the context lives there
    int prec;           // 4 or 8
    size_t esz;
    omx::Layout L;
    hipStream_t s = nullptr;
    int rows_y = 1, xchunk = 4;
    int tile_order = 0;         // tile order of the H/E kernels (YeeArgs::tile_order)
    int offset_bits = 32;       // 32: byte offsets where every array is below 4 GiB; 64: element indices

    void* f[6] = {};    // Ex Ey Ez Hx Hy Hz (current)
    void* f_alt[6] = {};        // second buffers of the fused step
    void* psi_h_alt[6] = {};
    int step_mode = 1;          // 0: fused step where it applies, 1: separate H and E kernels
    int wg_rows = 2;            // fused step: rows per wavefront (1 or 2)
    int wg_xchunk = 32;         // fused step: planes per workgroup tile
    int wg_ny = 0;              // fused step: wavefronts along y per workgroup (0: chosen from the z extent)
    int wg_med = -1;            // fused step needs its media variant: -1 unknown, 0 no, 1 yes
    int src_in_dispersive = -1; // -1 unknown, 0 no, 1 a source point sits in a Drude/DL component
    bool last_fused = false;
    // FP64 master coefficients: be ce (npml+1) bh ch (npml) den_e{x,y,z} den_h{x,y,z} (n+2 each)
    double* dcoef = nullptr;
    void* tcoef = nullptr;          // precision-T copies used by the kernels
    long co_be, co_ce, co_bh, co_ch, co_den[6], co_cd[6];
    long ncoef = 0;
    unsigned cpml_mask = 0;
    int nsh = 0, nse = 0;           // psi slots per CPML axis (H, E)
    void* psi_h[6] = {};
    void* psi_e[6] = {};

    // media
    bool media_set = false;
    int n_media = 0;
    omx_medium media[OMX_MAX_MEDIA + 1];
    int8_t* d_aux = nullptr;        // nx*ny*nz, Fortran order
    uint8_t* d_tag = nullptr;
    int8_t* d_cls = nullptr;        // 3 * n_classes
    int n_classes = 0;
    int disp_tag0 = 256;            // tags >= disp_tag0: classes with a Drude / DL component
    long* d_disp = nullptr;         // cells of those classes (memory order): i | j << 16 | k << 32 | media << 48
    int n_disp = 0;
    void* d_med = nullptr;          // MedDev<T>[n_media+1]
    omx_medium* d_med_in = nullptr;
    omx_medium_coeffs* d_med_co = nullptr;
    double* d_eps_class = nullptr;  // eps_r per class and component (energy)
    int maxp = 0;
    bool has_drude = false, has_dl = false;
    // Drude / DL state, compact: entry p belongs to the dispersive list's
    // cell p; per component E_old, PD, maxp PL and maxp PL_old planes of
    // n_disp entries (YeeArgs::dstate)
    void* disp_state = nullptr;

    // sources
    std::vector<omx::SourceDev> sources;

    // detectors
    std::vector<omx::DetectorHost> dets;
    long n_samples = 0;
    int* d_det_field = nullptr;     // per sample: field code*2 + point flag
    long* d_det_off = nullptr;
    int det_cap_samples = 0;
    double* d_rec = nullptr;
    int rec_cap = 0, rec_rows = 0;
    bool det_dirty = false;
    std::vector<int> h_det_field;
    std::vector<long> h_det_off;

    // quantum sites ("material" group)
    int nq = 0, nslot = 0, q_groups = 0;
    double q_density = 0.0, q_dt = 0.0;
    // quantum sub-stepping of the driver loop (mxim_mxll.f90:73-104): the
    // systems move every n_skip steps, q_dt = n_skip * dt; tq_step_old and
    // t_skip carry over from one omx3d_run call to the next
    int n_skip = 1, tq_step_old = 0;
    double t_skip = 0.0;
    long* d_q_off = nullptr;        // per site: cell offset
    omx::QSlot* d_q_slot = nullptr;      // per slot: its writers (k_scatter_j)
    double* d_J = nullptr;          // per slot: Jx Jy Jz Jold(3) dJ(3) -> 9*nslot
    double* d_q_e = nullptr;        // per site E (3)
    double* d_q_dpdt = nullptr;     // per site dPt/dt (3)
    int n_aff[3] = {0, 0, 0};
    long* d_aff_off[3] = {};
    int* d_aff_a[3] = {};
    int* d_aff_b[3] = {};
    int* d_kinds = nullptr;         // medium kind per medium index (gate of the J term)
    bool tls = false;
    double* d_tls = nullptr;        // per site: omega0 mu pol(3) gamma2 | rho(4) dip(3) dip_old(3)

    // near-field patches (FP32 grids with point sources): FP64 shadows of E
    // and H in boxes of patch_radius cells around the sources
    int patch_radius = 16;
    bool patches_built = false;     // boxes match the current sources / sites / media
    bool patch_dirty = false;       // shadows must be read again from the fields
    std::vector<omx::PatchBox> h_boxes;
    omx::PatchBox* d_boxes = nullptr;
    long patch_cells = 0;
    long patch_bytes = 0;           // device bytes the patches hold
    double* d_patch = nullptr;      // E shadow (3 * cells) then H shadow (3 * cells)
    double* d_patch_cdt = nullptr;  // dt_eps0 / eps_r per class and component
    int* d_patch_dl = nullptr;      // per box cell: position in the dispersive list or -1
    double* d_patch_ds = nullptr;   // FP64 Drude / DL state of the box cells
    std::vector<int> h_aff_ijk[3];  // cells whose E component sees J (3 ints each)
    double** d_aff_sh[3] = {};      // their shadow elements (null: none)

    // TF/SF
    bool tfsf = false;
    int tf_lo[3], tf_hi[3];
    int tf_kb = 0, tf_n = 0, tf_ks = 0;
    double tf_e0, tf_w0, tf_t0, tf_tau;
    double* d_tf = nullptr;         // ex1 (n) hy1 (n)
    int tf_fold_e = -1;             // TF/SF E corrections inside the E tiles: -1 unknown, 0 no, 1 yes

    // energy scratch
    double* d_red = nullptr;
    int red_blocks = 0;

    // profiling
    int prof_every = 0;     // events around the H/E launches of every prof_every-th step (0: none)
    std::vector<hipEvent_t> ev;
    double prof_ms_h = 0.0, prof_ms_e = 0.0;
    int prof_n = 0;

    long device_bytes = 0;
};
