/*
 * omx_fdtd.h -- C-ABI of libomx_fdtd_hip.so, the MI355X (gfx950) engine for the
 * 3D field-propagation hot path of One-Mxll-to-Rule-Them-All (OMxRTA).
 *
 * Every entry point takes plain pointers and sizes, returns an int status
 * (OMX_OK == 0) and never throws.  The text of the last failure on the calling
 * thread is returned by omx_last_error().  Host arrays that mirror reference
 * arrays use the reference's Fortran layout (first index fastest) and the
 * reference's index bounds; the device layout is private (z fastest, padded).
 *
 * Reference interfaces replaced (file:line relative to the reference tree):
 *   TMxll%init           src/FDTD/mxll_base_mod.f90:27,36-51  -> omx3d_create (+ set_media)
 *   TMxll%kill           src/FDTD/mxll_base_mod.f90:28,53-56  -> omx3d_destroy
 *   TMxll%td_propagate_H src/FDTD/mxll_base_mod.f90:29,58-61  -> omx3d_step_h
 *   TMxll%td_propagate_E src/FDTD/mxll_base_mod.f90:30,63-67  -> omx3d_step_e
 * Callers of the hot path that touch the solver's arrays directly:
 *   source_interactions (3D)  src/FDTD/interactions_mod.f90:102-156      -> omx3d_add_source / omx3d_inject_sources
 *   send_E_3D_to_J_ranks      src/FDTD/interactions_mod.f90:534,654-691  -> omx3d_gather_e
 *   send_J_to_E_3D_ranks      src/FDTD/interactions_mod.f90:1046,1177-1218 -> omx3d_scatter_j
 *   write_3D_field (stencils) src/FDTD/write_fields_subs_mod.f90:751-1509 -> omx3d_add_detector / omx3d_record
 */
#ifndef OMX_FDTD_H
#define OMX_FDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define OMX_OK          0
#define OMX_ERR_ARG     1   /* invalid argument or shape                    */
#define OMX_ERR_HIP     2   /* a HIP runtime call failed                     */
#define OMX_ERR_NOMEM   3   /* device allocation failed                      */
#define OMX_ERR_STATE   4   /* call out of order (e.g. media set twice)      */
#define OMX_ERR_RANGE   5   /* index or size beyond what the engine supports */

/* ---- codes shared with the reference (src/FDTD/constants_mod.f90:58-102) - */
#define OMX_BC_CLOSE     1
#define OMX_BC_PERIODIC  2
#define OMX_BC_CPML      3

#define OMX_MEDIUM_NONE        0
#define OMX_MEDIUM_DIELECTRIC  1
#define OMX_MEDIUM_DRUDE       2
#define OMX_MEDIUM_DL          3

/* q-group types (constants_mod.f90:74-75) */
#define OMX_Q_MATERIAL  1
#define OMX_Q_SINGLE    2

#define OMX_DET_POINT     1
#define OMX_DET_LINE_X    2
#define OMX_DET_LINE_Y    3
#define OMX_DET_LINE_Z    4
#define OMX_DET_PLANE_XY  5
#define OMX_DET_PLANE_YZ  6
#define OMX_DET_PLANE_ZX  7
#define OMX_DET_VOLUME    8

#define OMX_EX 1
#define OMX_EY 2
#define OMX_EZ 3
#define OMX_HX 4
#define OMX_HY 5
#define OMX_HZ 6

/* ---- engine limits ------------------------------------------------------ */
#define OMX_MAX_POLES   8    /* Lorentz poles per Drude-Lorentz medium      */
#define OMX_MAX_MEDIA   15   /* media per grid (cell tags are one byte)     */

#define OMX_FP32 4
#define OMX_FP64 8

typedef struct omx3d_ctx omx3d_ctx;

/* Grid description: the arguments of TMxll%init (mxll_base_mod.f90:36-51)
 * for a single, non-MPI domain. */
typedef struct omx3d_desc {
    int32_t nx, ny, nz;        /* grid_Ndims                                  */
    int32_t npml;              /* CPML thickness in cells (>= 2 if any CPML)  */
    int32_t boundaries[3];     /* OMX_BC_* per axis                           */
    int32_t precision;         /* OMX_FP32 or OMX_FP64 field arithmetic       */
    double  dt, dr;            /* atomic units                                */
} omx3d_desc;

/* One classical medium (TClassicalMedium, classical_medium_mod.f90:6-37).
 * Frequencies and rates in atomic units.  a_k = f_k * omega_p^2. */
typedef struct omx_medium {
    int32_t kind;              /* OMX_MEDIUM_*                                */
    int32_t n_poles;           /* Lorentz poles (DL only)                     */
    double  eps_r;             /* dielectric eps_r, or eps_inf of Drude / DL  */
    double  omega_d, gamma_d;  /* Drude pole                                  */
    double  omega_k[OMX_MAX_POLES];
    double  gamma_k[OMX_MAX_POLES];
    double  a_k[OMX_MAX_POLES];
} omx_medium;

/* Update coefficients of one medium, computed on the device in FP64 with the
 * formulas of init_classical_medium (classical_medium_mod.f90:60-101). */
typedef struct omx_medium_coeffs {
    double A1, A2, C1, C2, C3, C4;
    double inv_eps;            /* 1/eps_r (dielectric)                        */
    double alpha_k[OMX_MAX_POLES], beta_k[OMX_MAX_POLES], gamma_k[OMX_MAX_POLES];
    double B1_k[OMX_MAX_POLES], B2_k[OMX_MAX_POLES];
} omx_medium_coeffs;

/* Geometric primitive rasterised on the device into the cell medium map.
 * Coordinates are in cell-index units: cell (i,j,k) (1-based) sits at (i,j,k).
 * medium == 0 carves (sets cells back to no medium); otherwise a cell takes the
 * first medium that covers it, as the reference's point-list reader does
 * (classical_medium_mod.f90:667-686). */
#define OMX_SHAPE_SPHERE      1   /* p: cx cy cz r                              */
#define OMX_SHAPE_ELLIPSOID   2   /* p: cx cy cz a b c R00 R01 .. R22 (row-major body<-grid rotation) */
#define OMX_SHAPE_BOX         3   /* p: xlo ylo zlo xhi yhi zhi                 */
#define OMX_SHAPE_CYLINDER_Z  4   /* p: cx cy r zlo zhi                         */
typedef struct omx_shape {
    int32_t type;
    int32_t medium;
    double  p[15];
} omx_shape;

const char* omx_last_error(void);
int omx_version(void);
/* Library pole data of the reference (get_DL_library_data,
 * classical_medium_mod.f90:125-199), including its single-precision literals.
 * name: "Ag", "Al" or "Au". */
int omx_medium_library(const char* name, double eps_r, omx_medium* out);

/* ---- lifecycle (TMxll%init / TMxll%kill) --------------------------------- */
int omx3d_create(const omx3d_desc* desc, omx3d_ctx** out);
int omx3d_destroy(omx3d_ctx* ctx);
/* Tile shape of the H/E kernels: y rows per workgroup (1, 2 or 4), x planes
 * marched per workgroup (>= 1; 0 keeps the current values) and the order in
 * which workgroups take tiles: 0 block order (consecutive workgroups,
 * spread over the XCDs, cover consecutive memory), 1 one contiguous slab of
 * tiles per XCD, 2 block order by whole row blocks (the z tiles of a row
 * block on one XCD) (-1 keeps it).  A batch uses the tile order of its first
 * grid.  Defaults: FP32 4 rows x 4 planes, FP64 2 rows x 4 planes, tile
 * order 2.  No reference counterpart (engine tuning). */
int omx3d_set_tiling(omx3d_ctx* ctx, int32_t rows_y, int32_t xchunk, int32_t tile_order);
int omx3d_get_tiling(omx3d_ctx* ctx, int32_t* rows_y, int32_t* xchunk, int32_t* tile_order);
/* Addressing of the step kernels: 32 (default) uses 32-bit byte offsets when
 * every array of the grid is below 4 GiB, 64-bit element indices otherwise;
 * 64 always uses 64-bit element indices (same results).  No reference
 * counterpart (engine tuning). */
int omx3d_set_addressing(omx3d_ctx* ctx, int32_t offset_bits);
/* CPML profiles and den arrays (mxll_3D_mod.f90:282-420), device FP64.
 * be,ce: npml; bh,ch: npml-1; den: den_ex(nx) den_ey(ny) den_ez(nz) den_hx(nx) den_hy(ny) den_hz(nz). */
int omx3d_get_cpml(omx3d_ctx* ctx, double* be, double* ce, double* bh, double* ch, double* den);

/* ---- media (read_init_media_3D, classical_medium_mod.f90:546-769) -------- */
/* aux: the cell medium map (nx*ny*nz, Fortran order, values 0..n_media) the
 * reference builds from the medium point lists; the per-component staggered
 * map is derived from it on the device. */
int omx3d_set_media(omx3d_ctx* ctx, int32_t n_media, const omx_medium* table, const int8_t* aux);
int omx3d_set_media_shapes(omx3d_ctx* ctx, int32_t n_media, const omx_medium* table,
                           int32_t n_shapes, const omx_shape* shapes);
/* media_map(nx,ny,nz,3) of the reference, Fortran order. */
int omx3d_get_media_map(omx3d_ctx* ctx, int32_t* media_map);
int omx3d_get_medium_coeffs(omx3d_ctx* ctx, int32_t medium, omx_medium_coeffs* out);

/* ---- time stepping (td_propagate_H_3Dfield / td_propagate_E_3Dfield) ----- */
int omx3d_step_h(omx3d_ctx* ctx);
/* time and t_skip are TMxll%time and TMxll%t_skip (mxim_mxll.f90:93-101). */
int omx3d_step_e(omx3d_ctx* ctx, double time, double t_skip);

/* ---- sources (TSource + source_interactions) ----------------------------- */
/* ind: (2n+1)^3 x 3 Fortran array (ind_i, ind_j, ind_k blocks, i fastest);
 * ker: ker_mat(-n:n,-n:n,-n:n).  dir: 1,2,3 for x,y,z.  Times in au. */
int omx3d_add_source(omx3d_ctx* ctx, int32_t dir, double j_amp, double w0, double t0, double tau,
                     double t_init, double t_final, double phase, int32_t n_ker,
                     const int32_t* ind, const double* ker);
int omx3d_inject_sources(omx3d_ctx* ctx, double time);

/* ---- detectors (TDetector + write_3D_field stencils) -------------------- */
/* indx_list(nd,3), Fortran order, 1-based local indices. */
int omx3d_add_detector(omx3d_ctx* ctx, int32_t det_type, int32_t field, int32_t nd,
                       const int32_t* indx_list);
/* Device ring of samples: capacity rows of (sum of nd) doubles. */
int omx3d_set_record_capacity(omx3d_ctx* ctx, int32_t rows);
int omx3d_record(omx3d_ctx* ctx);
int omx3d_read_records(omx3d_ctx* ctx, double* out, int32_t* rows);
int omx3d_sample(omx3d_ctx* ctx, double* out);

/* ---- embedded quantum systems (q-groups) -------------------------------- */
/* n quantum systems in n_groups q-groups (TQ_group, q_group_mod.f90), the
 * systems listed in group order (group_of[q], 0-based, non-decreasing; may be
 * null for one group).  ijk: the cell of each system as an (n,3) Fortran
 * array -- map(n,3:5) of a "material" group (q_group_mod.f90:505-509), the
 * kernel centre kernel_map(n,0,0,0,3:5) of a "single" one (:717-734).  E is
 * gathered at that cell (interactions_mod.f90:654-691).  J: a material
 * system writes density*dPt/dt into its cell, a single one
 * density*w*dPt/dt into the 7^3 cells around it with the unnormalised
 * Blackman-Harris weights w of q_group_mod.f90:216-239; the last writer of a
 * cell wins and, as the reference copies J_old <- J at every group's
 * scatter, only the last group's writes leave a dJ
 * (interactions_mod.f90:1046-1218).  Every kernel cell must lie in the grid.
 * dt_q (au): the requested quantum step.  As the driver loop does
 * (mxim_mxll.f90:73-79), the systems move every n_skip = NINT(dt_q/dt) steps
 * when dt_q > dt (else every step) and are propagated over n_skip*dt;
 * omx3d_run sets t_skip at each new quantum step (mxim_mxll.f90:97-104). */
int omx3d_set_q_groups(omx3d_ctx* ctx, int32_t n_groups, const int32_t* group_type, int32_t n,
                       const int32_t* group_of, const int32_t* ijk, double density, double dt_q);
/* One "material" group of n systems. */
int omx3d_set_q_sites(omx3d_ctx* ctx, int32_t n, const int32_t* ijk, double density, double dt_q);
int omx3d_gather_e(omx3d_ctx* ctx, double* e_sites);              /* (n,3) */
int omx3d_scatter_j(omx3d_ctx* ctx, const double* dpdt_sites);    /* (n,3) */
/* Two-level emitters propagated on the device in FP64 (RK4 over dt_q). */
int omx3d_set_two_level(omx3d_ctx* ctx, const double* omega0, const double* mu,
                        const double* pol, const double* gamma2);
int omx3d_two_level_step(omx3d_ctx* ctx);                          /* gather, RK4, scatter */
int omx3d_get_two_level(omx3d_ctx* ctx, double* dipole, double* rho); /* (n,3), (n,4) */

/* ---- field access and reductions ---------------------------------------- */
/* host: the reference array of that field with its bounds (mxll_3D_mod.f90:123-131),
 * Fortran order.  to_device = 0 copies device -> host. */
int omx3d_copy_field(omx3d_ctx* ctx, int32_t field, double* host, int32_t to_device);
/* W = 1/2 dr^3 sum(eps0 eps_r E^2 + mu0 H^2) over the interior cells; w64 is
 * accumulated in FP64, w32c with compensated FP32 sums. */
int omx3d_energy(omx3d_ctx* ctx, double* w64, double* w32c);
/* omx3d_energy for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops): no reference counterpart.  Allocates
 * nothing: the reduction's partial sums go to a fixed 24 KiB device array of
 * the library (one per GPU, for the life of the process), and (w64, w32c)
 * are written to w_dev, 2 doubles of device memory on the context's GPU,
 * asynchronously on the context's stream (omx3d_synchronize before reading).
 * Same kernels and launch shape as omx3d_energy: the same values, bit for
 * bit.  The scratch does not grow with the grid (the reduction strides over
 * every cell with a fixed number of workgroups).  Calls that overlap on one
 * GPU share that array and race: issue them from one stream at a time. */
int omx3d_energy_noalloc(omx3d_ctx* ctx, double* w_dev);

/* ---- total-field / scattered-field plane wave (x-polarised, along +z) ---- */
int omx3d_set_tfsf(omx3d_ctx* ctx, const int32_t lo[3], const int32_t hi[3],
                   double e0, double w0, double t0, double tau, int32_t max_steps);
/* The plane wave's incident line (after the steps queued so far): Ex at
 * k = kb .. kb+n-1 and Hy at k+1/2, n values each (ex1 / hy1 may be null to
 * query kb and n).  No reference counterpart: it lets a caller compare the
 * total field inside the box with the incident wave. */
int omx3d_get_tfsf_line(omx3d_ctx* ctx, double* ex1, double* hy1, int32_t* kb, int32_t* n);

/* ---- whole time loop on the device (mxim_mxll.f90:92-132 step order) ----- */
/* Steps tt = tt_first .. tt_first+n_steps-1; records detectors when
 * mod(tt, record_every) == 0 (record_every <= 0: never). */
int omx3d_run(omx3d_ctx* ctx, int32_t tt_first, int32_t n_steps, int32_t record_every);
/* Step kernels of omx3d_run: mode 1 (default) runs the separate H and E
 * kernels; mode 0 runs H and E as one fused pass over HBM (fields
 * double-buffered) wherever it applies -- no periodic axis, no source point
 * in a Drude/DL component, and a z extent one workgroup covers (nz <= 1024
 * in FP32, 512 in FP64).  A fused workgroup holds every z cell of a block of
 * rows and marches a chunk of planes; rows (1 or 2 rows of 64 z cells per
 * wavefront) and xchunk (planes per workgroup) set its tile; 0 keeps the
 * current value.  No reference counterpart. */
int omx3d_set_step_mode(omx3d_ctx* ctx, int32_t mode, int32_t rows, int32_t xchunk);
/* Wavefronts along y of a fused workgroup (1..16, at most 16 (FP32) or 8
 * (FP64) wavefronts in all); 0 chooses them from the z extent.  No
 * reference counterpart. */
int omx3d_set_fused_waves(omx3d_ctx* ctx, int32_t ny_waves);
/* The fused tile in use: rows per wavefront, planes per chunk, wavefronts
 * along y (as resolved for this grid; 0 when the z extent does not fit).
 * No reference counterpart. */
int omx3d_get_fused_shape(omx3d_ctx* ctx, int32_t* rows, int32_t* xchunk, int32_t* ny_waves);
/* 1 if the last omx3d_run used the fused step. */
int omx3d_get_step_mode(omx3d_ctx* ctx, int32_t* fused_last_run);
/* Per-kernel timing: HIP events around the H and E launches of every
 * every-th step of omx3d_run (1: every step, 0: off); omx3d_get_profile
 * sums them.  With the fused step, ms_h is 0 and ms_e times the fused
 * kernel.  Each event pair costs the stream a few microseconds, so a
 * stride above 1 keeps the timed run close to an unprofiled one. */
int omx3d_set_profiling(omx3d_ctx* ctx, int32_t every);
int omx3d_get_profile(omx3d_ctx* ctx, double* ms_h, double* ms_e, int32_t* n_launches);
/* Near-field patches of FP32 grids.  The quasi-static near field of a
 * point source is (c/(omega dr))^2 times its radiated field one cell away;
 * rounding it to FP32 every half step adds noise that reaches the far
 * detectors at ~1e-4 of their signal.  An FP32 grid therefore keeps FP64
 * copies of E and H in a box of `radius` cells around each source (boxes
 * outside the CPML, merged where they touch), recomputes the box in FP64
 * after each H and E update and writes the rounded result to the fields;
 * sources and quantum-site currents inside a box act on the FP64 copy.
 * Default radius 16 (at dx = 1 nm and a 600 nm carrier, the FP32 error of
 * far detectors over 10000 steps is 1.6e-5 with radius 8 and 6e-6 with 16:
 * rounding the near field while the source runs leaves small spurious
 * charges whose static field persists); 0 turns the patches off; FP64
 * grids and grids with a
 * TF/SF plane wave have none.  omx3d_get_source_patch reports the radius and
 * the boxes and cells in use.  No reference counterpart (FP32 accuracy). */
int omx3d_set_source_patch(omx3d_ctx* ctx, int32_t radius);
int omx3d_get_source_patch(omx3d_ctx* ctx, int32_t* radius, int32_t* n_boxes, int64_t* n_cells);
int omx3d_synchronize(omx3d_ctx* ctx);
/* Bytes of device memory the context holds. */
int omx3d_device_bytes(omx3d_ctx* ctx, int64_t* bytes);
/* Tile-shape autotuner: times rows_y x xchunk x tile order candidates with
 * `steps` H+E steps each, keeps the fastest, then clears fields, CPML and
 * dispersive state (call before a run).  *rows_y on entry: 0 tunes the row
 * count too; 1, 2 or 4 keeps it (the grids of a batch share one workgroup
 * height).  No reference counterpart (engine tuning). */
int omx3d_autotune(omx3d_ctx* ctx, int32_t steps, int32_t* rows_y, int32_t* xchunk, int32_t* tile_order,
                   double* ms_per_step);
/* Step-kernel autotuner: times `steps` steps of the separate H and E
 * kernels (at the current tiling) and, where the fused step applies, of
 * fused workgroup tiles of several shapes, on the grid itself (TF/SF and
 * near-field patches included, sources and detectors left out); keeps the
 * fused step (and its fastest shape) when it beats the separate kernels by
 * more than 2 %, else the separate kernels (omx3d_set_step_mode); frees the fused
 * step's second buffers when the separate kernels win, then clears fields,
 * CPML, dispersive state and the TF/SF line (call before a run).  *mode: 0
 * fused, 1 separate; *ms_fused < 0 when the fused step does not apply.  No
 * reference counterpart (engine tuning). */
int omx3d_autotune_step(omx3d_ctx* ctx, int32_t steps, int32_t* mode, double* ms_separate, double* ms_fused);

/* ---- batches of independent grids (no reference counterpart) ------------ */
/* One launch per phase (H, sources, E, detectors) covers every grid of the
 * batch; each grid computes exactly what its own omx3d_run computes.  The
 * grids are set up with the single-grid calls first (same precision, no
 * TF/SF or quantum sites); use_graph replays a captured step graph.  The
 * grids stay owned by the caller and must outlive the batch; omxb_run picks
 * up changes made to them since the last call (buffers, sources, detectors,
 * tiling, record buffer).  Recording into a full detector buffer fails as in
 * omx3d_run. */
typedef struct omxb_ctx omxb_ctx;
int omxb_create(int32_t n, omx3d_ctx* const* grids, omxb_ctx** out);
int omxb_run(omxb_ctx* batch, int32_t tt_first, int32_t n_steps, int32_t record_every, int32_t use_graph);
int omxb_destroy(omxb_ctx* batch);

/* ---- 2D and 1D grids (TMxll_2D, mxll_2D_mod.f90; TMxll_1D, mxll_1D_mod.f90) */
/* FP64 only.  2D: n1 = nx, n2 = ny, mode OMX_2D_TMZ (Ez, Hx, Hy), OMX_2D_TEZ
 * (Ex, Ey, Hz) or OMX_2D_FULL (both families, which do not couple in 2D),
 * boundaries[0..1] per axis.  1D (the reference's z line: Ex at k, Hy at
 * k+1/2): n1 = nz, n2 and mode ignored, boundaries[0].  Arrays keep the
 * reference's bounds (Ex(0:nx,1:ny+1), Ey(1:nx+1,0:ny), Ez(1:nx+1,1:ny+1),
 * Hx(1:nx,0:ny), Hy(0:nx,1:ny), Hz(0:nx,0:ny); 1D Ex(1:nz), Hy(1:nz-1)),
 * which omxl_copy_field reads and writes in Fortran order.  The reference's
 * FULL mode allocates only the TEz arrays (mxll_2D_mod.f90:110-163) and
 * cannot run; here it runs both families. */
#define OMX_2D_TMZ  1
#define OMX_2D_TEZ  2
#define OMX_2D_FULL 3
typedef struct omxl_desc {
    int32_t dims;              /* 1 or 2                                      */
    int32_t mode;              /* OMX_2D_* (2D)                               */
    int32_t n1, n2;            /* cells: nx, ny (2D) or nz (1D)               */
    int32_t npml;              /* CPML thickness (>= 2 if any CPML)           */
    int32_t boundaries[2];     /* OMX_BC_* per axis                           */
    int32_t precision;         /* OMX_FP64                                    */
    double  dt, dr;            /* atomic units                                */
} omxl_desc;
typedef struct omxl_ctx omxl_ctx;
/* init_2Dgrid / init_1Dgrid (mxll_2D_mod.f90:177-420, mxll_1D_mod.f90:52-210):
 * fields zero, CPML profiles and den arrays on the device. */
int omxl_create(const omxl_desc* desc, omxl_ctx** out);
int omxl_destroy(omxl_ctx* ctx);
/* read_init_media_2D / read_init_media_1D (classical_medium_mod.f90:341-538,
 * 203-337): aux is the cell medium map (nx*ny Fortran order, or nz), values
 * 0..n_media.  2D keeps the reference's staggering: component maps from the
 * x / y neighbours, the edge row and column set only in the Ex map, and
 * every component's eps from the Ex map's dielectric. */
int omxl_set_media(omxl_ctx* ctx, int32_t n_media, const omx_medium* table, const int8_t* aux);
/* td_propagate_H_2Dfield / _1Dfield and td_propagate_E_* (mxll_2D_mod.f90:450-864,
 * mxll_1D_mod.f90:252-352), with the periodic copies of their non-MPI branches. */
int omxl_step_h(omxl_ctx* ctx);
int omxl_step_e(omxl_ctx* ctx);
/* A source of read_init_sources (source_mod.f90:40-100): dir 1..3 (x, y, z;
 * 1D: 1), ind(w*w, 2) (2D, Fortran order; 1D: ind(w)) 1-based indices and
 * ker the kernel, w = 2 n_ker + 1.  source_interactions, 1D / 2D
 * (interactions_mod.f90:32-100): E += dt_eps0/dr/c0/2 * J, with J the mean of
 * the point and its +x (+y) neighbour in the kernel for 2D Ex (Ey), the
 * point's own J for 2D Ez and 1D Ex. */
int omxl_add_source(omxl_ctx* ctx, int32_t dir, double j_amp, double w0, double t0, double tau, double t_init,
                    double t_final, double phase, int32_t n_ker, const int32_t* ind, const double* ker);
int omxl_inject_sources(omxl_ctx* ctx, double time);
/* Samples of field at nd cells idx(nd, dims) (Fortran order, 1-based) with the
 * stencils of write_2D_field / write_1D_field (write_fields_subs_mod.f90:88-628). */
int omxl_add_detector(omxl_ctx* ctx, int32_t field, int32_t nd, const int32_t* idx);
int omxl_sample(omxl_ctx* ctx, double* out);
/* Steps tt = tt_first .. tt_first+n_steps-1 in the driver loop's order
 * (mxim_mxll.f90:92-132); after every step with mod(tt, record_every) == 0
 * the samples of every detector go to the next row of out. */
int omxl_run(omxl_ctx* ctx, int32_t tt_first, int32_t n_steps, int32_t record_every, double* out);
int omxl_copy_field(omxl_ctx* ctx, int32_t field, double* host, int32_t to_device);
int omxl_synchronize(omxl_ctx* ctx);
int omxl_device_bytes(omxl_ctx* ctx, int64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* OMX_FDTD_H */
