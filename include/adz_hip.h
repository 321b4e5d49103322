/* adz_hip.h -- C ABI of libadz_hip.so, the MI355X (gfx950) implementation of GEM's
 * semi-Lagrangian advection step (gemdyn adz_* layer).
 *
 * Every entry point is extern "C" with plain pointers and sizes; device arrays are
 * hipMalloc'd by the caller unless stated otherwise.  Every function returns 0 on success
 * or an ADZ_E* code; adz_last_error() gives the message of the last failure on the calling
 * thread.  Nothing here throws across the boundary.
 *
 * Array conventions follow the reference: Fortran column-major (i fastest), positions in
 * GLOBAL index space with f(1,1,1) at (1.0,1.0,1.0), AoS (px,py,pz) triplets, and source
 * pointers that address f(1,1,1) of each (possibly halo-extended) array.
 */
#ifndef ADZ_HIP_H
#define ADZ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ADZ_OK 0
#define ADZ_EINVAL 1 /* bad argument (null pointer, size, dtype)            */
#define ADZ_EHIP 2   /* a HIP runtime call failed                           */
#define ADZ_ENOMEM 3 /* device allocation failed                            */
#define ADZ_ESTATE 4 /* call out of order (e.g. tracers before trajectory)  */

/* storage types of fields and positions */
#define ADZ_F64 0
#define ADZ_F32 1
#define ADZ_BF16 2

/* arithmetic of the interpolation */
#define ADZ_MATH_REF 0 /* FP64 weights and accumulation in the reference's x86 AVX2+FMA
                          operation order (turbo_tricublin.c:553-615); bit-exact with the
                          reference build for FP32 fields                              */
#define ADZ_MATH_F32 1 /* index/floor step in FP64, weights and accumulation in FP32    */

typedef struct adz_ztab adz_ztab; /* device-resident vertical descriptor (opaque) */
typedef void* adz_stream;         /* a hipStream_t; NULL = the null stream         */

const char* adz_last_error(void);
int adz_version(void);

/* ------------------------------------------------------------------------------------
 * Level 1: the interpolation kernel (replaces turbo_tricublin.c / tricublin_f90.inc).
 * ----------------------------------------------------------------------------------*/

/* Vertical descriptor.  Replaces Vsearch_setup_plus (turbo_tricublin.c:368-380, which
 * wraps Vsearch_setup, :320-363); `levels` is a HOST array of nk doubles.  Unlike the
 * reference (which never frees its malloc'd ztab) the descriptor is released by
 * adz_vsearch_free.  nk >= 4 is required (the reference silently fills 1.0 below that). */
int adz_vsearch_setup_plus(const double* levels, int nk, int ni, int nj, int offseti, int offsetj,
                           adz_ztab** out);
/* Allocated LOCAL index range of the source arrays addressed through this descriptor,
 * e.g. (1-halo, ni_local+halo) when f(1,1,1) sits inside a halo-extended array.  Default:
 * [1, ni] x [1, nj].  Targets whose 4x4 stencil leaves this range (the reference reads
 * out of bounds there, :744-747) produce NaN instead of a memory fault.  No reference
 * counterpart. */
int adz_vsearch_set_bounds(adz_ztab* lv, int ilo, int ihi, int jlo, int jhi);
int adz_vsearch_free(adz_ztab* lv);

/* Index and weights of n targets: Vcoef_pxyz4_inline (turbo_tricublin.c:382-512).
 * Device outputs: ijk[n] (offset of the stencil corner from f(1,1,1)), zlin[n] (the
 * zlinear flag), coef[20*n] (FP64 weights; may be NULL).  For parity checks. */
int adz_vcoef(int* ijk, int* zlin, double* coef, const void* pxyz, int pdtype, const adz_ztab* lv, int n,
              adz_stream stream);

/* Tricublin_zyx1_m_n (turbo_tricublin.c:793-815): d[f*n + p] = field f at target p.
 * fs: HOST array of m device pointers to f(1,1,1) of each source field.  d, fs[] have
 * storage type fdtype; pxyz has pdtype (ADZ_F32 = the reference's pxpypz). */
int adz_tricublin_zyx1_m_n(void* d, const void* const* fs, const void* pxyz, const adz_ztab* lv, int n, int m,
                           int fdtype, int pdtype, int math, adz_stream stream);

/* Tricublin_mono_zyx_m_n (turbo_tricublin.c:1080-1109): cubic result d, linear result l
 * and min / max of the inner cells (ADZ_MATH_REF follows the reference's AVX2 footprint,
 * x 2 columns * y 4 rows * z 2 planes). */
int adz_tricublin_mono_zyx_m_n(void* d, void* l, void* mi, void* ma, const void* const* fs, const void* pxyz,
                               const adz_ztab* lv, int n, int m, int fdtype, int pdtype, int math,
                               adz_stream stream);

/* Tricublin_zyx3_n (turbo_tricublin.c:1298-1312): d(3,n) from the interleaved f3(3,...). */
int adz_tricublin_zyx3_n(void* d, const void* f3, const void* pxyz, const adz_ztab* lv, int n, int fdtype,
                         int pdtype, int math, adz_stream stream);


/* ------------------------------------------------------------------------------------
 * Level 2: the adz_* layer (replaces the Fortran routines of gemdyn/src/base/adz_*.F90
 * on device-resident state).  One context per model tile.
 * ----------------------------------------------------------------------------------*/

/* What adz_set reads from GEM's modules (glb_ld, ver, geomh, dcst, lam/grid options).
 * zm/zt/lat are HOST arrays: Ver_z_8%m(0:nk+1), Ver_z_8%t(0:nk+1), Ver_z_8%x(0:nk) and
 * the latitude (radians) of local rows lminy..lmaxy. */
typedef struct adz_grid {
  int ni, nj, nk;                    /* l_ni, l_nj, l_nk (= G_nk)                        */
  int i0g, j0g;                      /* l_i0, l_j0: global index of local (1,1)         */
  int halox, haloy;                  /* Adz_halox / Adz_haloy (= Grd_maxcfl + 1)         */
  int west, east, south, north;      /* tile touches the domain edge (1/0)               */
  int pil_w, pil_e, pil_s, pil_n;    /* Glb_pil_* (used where the tile touches the edge) */
  int gbpil_t;                       /* Lam_gbpil_T                                      */
  int bc_lam_flux;                   /* Adz_BC_LAM_flux                                  */
  int niu, njv;                      /* l_niu, l_njv                                     */
  double inv_hx, inv_hy;             /* geomh_inv_hx_8, geomh_inv_hy_8                   */
  double inv_rayt;                   /* Dcst_inv_rayt_8                                  */
  const double* zm;
  const double* zt;
  const double* zx;
  const double* lat;
  double zmin, zmax;                 /* Ver_zmin_8, Ver_zmax_8                           */
  double wmstar_nk;                  /* Ver_wmstar_8(nk)                                 */
  double idz_t_nkm1;                 /* Ver_idz_8%t(nk-1)                                */
} adz_grid;

typedef struct adz_ctx adz_ctx;

/* adz_set (adz_set.F90:16-311, single tile, ADZ_OD_L = .false.): bounds, clamps, level
 * search tables, vertical Lagrange weights, cos-latitude factors, the two vertical
 * descriptors and the device state.  pos_dtype selects the storage of the derived winds
 * and departure positions: ADZ_F32 = the reference's declared `real` arrays
 * (adz_mem.F90:78-84), bit-exact trajectories; ADZ_F64 = FP64 throughout.  itraj is
 * Adz_itraj (adz_options.F90:30, default 3). */
int adz_set(const adz_grid* g, int pos_dtype, int itraj, adz_ctx** out);
int adz_ctx_free(adz_ctx* c);
/* Derived integers (24): lminx lmaxx lminy lmaxy nit njt nij i0 in j0 jn i0u inu j0v jnv
 * i0b inb j0b jnb k0 k0t k0m pnz_m pnz_t; doubles (9): iminposx imaxposx iminposy
 * imaxposy sig ovdzm ovdzt vw5 vw6.  No reference counterpart (module variables there). */
int adz_ctx_info(const adz_ctx* c, int* ints, double* dbls);
/* Copy a derived table to the host: 0 search_m, 1 search_t (int); 2..9 adz_vw1m..4m,
 * adz_vw1t..4t (nk+1 doubles, index k); 10 cy (njt); 11 odelz_m; 12 odelz_t. */
int adz_ctx_table(const adz_ctx* c, int which, void* host_out);

/* State arrays owned by the context (device pointers): */
#define ADZ_A_PXYZM 0   /* Adz_pxyzm(3, ni, nj, k0m:nk)           pos_dtype */
#define ADZ_A_WPXYZ 1   /* Adz_wpxyz(-1:ni+2, -1:nj+2, nk, 3)      FP64      */
#define ADZ_A_UVW_DEP 2 /* Adz_uvw_dep(3, ni, nj, k0m:nk)          pos_dtype */
#define ADZ_A_PM 3      /* Adz_pm (3, i0:in,   j0:jn,   k0:nk)     pos_dtype */
#define ADZ_A_PMU 4     /* Adz_pmu(3, i0u:inu, j0:jn,   k0:nk)     pos_dtype */
#define ADZ_A_PMV 5     /* Adz_pmv(3, i0:in,   j0v:jnv, k0:nk)     pos_dtype */
#define ADZ_A_PT 6      /* Adz_pt (3, i0:in,   j0:jn,   k0t:nk)    pos_dtype */
#define ADZ_A_PB 7      /* Adz_pb (3, i0b:inb, j0b:jnb, 1:nk)      pos_dtype */
#define ADZ_A_UVW_D 8   /* Adz_uvw_d(3, lminx:lmaxx, lminy:lmaxy, nk) pos_dtype */
#define ADZ_A_WARR 9    /* Adz_uu_arr, Adz_vv_arr, Adz_ww_arr (ni, nj, nk) x3   */
int adz_ctx_array(adz_ctx* c, int which, void** dptr, long* nelem);
/* The vertical descriptors adz_set builds (Adz_cpntr_q / Adz_cpntr_t); owned by c. */
const adz_ztab* adz_ctx_ztab(const adz_ctx* c, int thermo);

/* Restart of the semi-Lagrangian state: the arrays the reference checkpoints for the
 * scheme, ADZ_PXYZM (Adz_pxyzm, the context's position storage) and ADZ_WPXYZ (Adz_wpxyz,
 * FP64), registered with GMM_FLAG_RSTR at adz_set.F90:343-360, plus the iteration count of
 * the next adz_traject.  adz_ctx_save_state copies them into a HOST buffer of
 * adz_ctx_state_bytes bytes (fixed header, then the arrays), adz_ctx_load_state into a
 * context built by adz_set for the same grid and position storage; both synchronise the
 * stream.  A run that saves after step N and continues from a fresh context that loads
 * the state reproduces the uninterrupted run bit for bit (the reference's restart test).
 * The BC_mass_deficit of each tracer, which the reference writes in its restart file
 * (wrrstrt.F90:49-63), is the caller's device array and is saved by the caller. */
int adz_ctx_state_bytes(const adz_ctx* c, size_t* bytes);
int adz_ctx_save_state(const adz_ctx* c, void* host, size_t bytes, adz_stream stream);
int adz_ctx_load_state(adz_ctx* c, const void* host, size_t bytes, adz_stream stream);

/* adz_inittraj (adz_inittraj.F90:16-42): arrival points as first guess, and the next
 * trajectory uses max(6, itraj) iterations. */
int adz_inittraj(adz_ctx* c, adz_stream stream);

/* adz_traject (adz_traject.F90:16-130) with adz_prepareWinds and adz_interp_traj: winds
 * are FP32 DEVICE arrays on the extended layout (lminx:lmaxx, lminy:lmaxy, nk): uu_ext,
 * vv_ext, ww_ext (Adz_uu/vv/ww_ext), ut0, vt0 (staggered winds at time 0, halo >= 2) and
 * zdt0 (zeta-dot at time 0).  schm_advec, bA_m, bA as Schm_advec, Cstv_bA_m_8, Cstv_bA_8.
 * Fills every position set of the context. */
int adz_traject(adz_ctx* c, double dt, int schm_advec, double bA_m, double bA, const float* uu_ext,
                const float* vv_ext, const float* ww_ext, const float* ut0, const float* vt0, const float* zdt0,
                adz_stream stream);

/* ------------------------------------------------------------------------------------
 * Level 3: fused multi-field interpolation (adz_tricub_hlt) and the batched, persistent
 * form over several subdomains.
 * ----------------------------------------------------------------------------------*/

/* One field of a group (the reference's Adz_pntr_stack entry, adz_mem.F90:66-72, plus the
 * Adz_post arrays of adz_mem.F90:92-98).  All are DEVICE pointers to element (1,1,1). */
typedef struct adz_field {
  const void* src; /* source; halo-extended layout of `geom` when ext != 0, else the
                      model layout (dminx:dmaxx, dminy:dmaxy, nk)                     */
  void* dst;       /* destination, model layout                                       */
  void* lin;       /* Adz_post(n)%lin (clipped to [min,max]); with ADZ_POST_STORE      */
  void* mn;        /* Adz_post(n)%min                                                 */
  void* mx;        /* Adz_post(n)%max                                                 */
} adz_field;

#define ADZ_POST_STORE 1 /* store lin/min/max (adz_interp_hlt.F90:146-163)              */
#define ADZ_POST_CLIP 2  /* clip dst to [min, max] (mono=1, adz_post_tr.F90:81-94)      */

/* adz_tricub_hlt (adz_interp_hlt.F90:28-187): interpolate nptr fields at the positions
 * xyz of the target box (i0:in, j0:jn, k0:nk) (i fastest, the layout of Adz_pm/pmu/pmv/pt/
 * pb) and write each result into dst(i,j,k).  `fld` is a HOST array of nptr entries.
 * ext = 0 interpolates straight from the model array (no copy into an extended work
 * array: on one tile every stencil lies inside it; a stencil that does not yields NaN). */
int adz_tricub_hlt(const adz_field* fld, int nptr, const void* xyz, int pdtype, const adz_ztab* geom, int i0,
                   int in, int j0, int jn, int k0, int nk, int dminx, int dmaxx, int dminy, int dmaxy, int ext,
                   int post, int fdtype, int math, adz_stream stream);

/* adz_tricub_hlt with F_QV_L = .true. (adz_interp_hlt.F90:59-66,113-115): bicubic in x and
 * y, quintic in z (cubic at kk1 = 2 and nk-2, linear in the end intervals), with the
 * linear result and min / max of the inner 2x2x2 box, as bicubHQV_lin_min_max_ntr and
 * bicubHQV_lin_min_max_zyx compute them (bicubHQV_lin_mono.F90:20-172,
 * bicubHQV_lin_min_max_zyx.F90:20-99).  Positions have the context's storage type and
 * index the thermodynamic levels of ctx (Adz_pt); fields are FP32 (`real`); ext and post as
 * for adz_tricub_hlt.  The reference's arithmetic: FP64 weights and sums, FP32 row
 * results, no contraction. */
int adz_tricub_hlt_qv(adz_ctx* ctx, const adz_field* fld, int nptr, const void* xyz, int i0, int in, int j0, int jn,
                      int k0, int dminx, int dmaxx, int dminy, int dmaxy, int ext, int post, adz_stream stream);

/* A batched work item: one subdomain (or position set) with its own descriptor,
 * positions and field group (the arguments of one adz_tricub_hlt call). */
typedef struct adz_item {
  const adz_ztab* geom;
  const void* xyz;
  int pdtype;
  int i0, in, j0, jn, k0, nk;
  int dminx, dmaxx, dminy, dmaxy;
  int ext;
  const adz_field* fld; /* HOST array of nptr fields (copied into the plan) */
  int nptr;
} adz_item;

/* Tile parameters of the persistent kernel; zero fields_per_pass / blocks_per_cu take the
   built-in default, {10, 0, 1, 1, 8, 1} for FP32 fields in the FP32 arithmetic and
   {10, 0, 4, 1, 5, 3} otherwise, and stage 0 takes the built-in staging together with its
   level_inner; adz_plan_autotune searches them on the actual plan. */
typedef struct adz_tune {
  int fields_per_pass; /* fields interpolated per pass over the points (1..64)          */
  int points_outer;    /* 1: a block runs every field group of a point tile; 0: a pass
                          over all points per field group                               */
  int blocks_per_cu;   /* persistent grid = blocks_per_cu * CUs                         */
  int xcd_chunks;      /* 1: each XCD takes a contiguous range of tiles                 */
  int stage;           /* 1: stage each tile's source box in LDS; 2: per-lane loads;
                          3 (FP32 sources): one producer wave per block streams the
                          boxes of several fields ahead into LDS (DMA), four consumer
                          waves evaluate from LDS; 4 (FP32 sources): each lane loads one
                          stencil column, reduces it for its own and its three left
                          neighbours' targets, partials shared by lane shifts (61-wide
                          tiles); 5 (FP32 sources): all four waves copy each field's
                          box into LDS with 16-byte DMA, double-buffered, and every
                          lane evaluates from LDS; 6 (FP32 sources): the same with
                          tiles of 64 x 8 targets and blocks of eight waves; 7 and 8
                          (FP32 sources): tiles of 64 x 4 targets over 2 and 4
                          consecutive levels, blocks of 8 and 16 waves (8: the FP32
                          arithmetic only, ADZ_EINVAL otherwise; one block per CU is
                          resident; the fastest form at c5)                             */
  int level_inner;     /* tile order. 0: i, j, then k; 1: consecutive tiles walk the
                          levels of one tile column; S >= 2: super-columns of S tile
                          rows walked level by level (source planes and rows reused
                          in L2), then i, then j                                         */
} adz_tune;

typedef struct adz_plan adz_plan;
/* Build a plan for nitems subdomains: uploads the descriptors once, so that executing
 * it launches ONE persistent kernel with no host work (capturable in a hipGraph).
 * No reference counterpart (the reference loops over slices on OpenMP threads). */
int adz_plan_create(const adz_item* items, int nitems, int fdtype, int math, int post, const adz_tune* tune,
                    adz_plan** out);
int adz_plan_execute(const adz_plan* p, adz_stream stream);
int adz_plan_tune(const adz_plan* p, adz_tune* out); /* the parameters in use */
/* Staging waits of the plan's launches (stage 3) that gave up since the last call, read
 * from a device counter and reset; the fields they concern were evaluated from global
 * memory by per-lane loads, so the results stay exact.  Synchronises the device.  No
 * reference counterpart. */
int adz_plan_status(adz_plan* p, int* aborted_waits);
/* Polls of one bounded staging wait before it gives up (default 2^20); 0 makes every wait
 * that is not already satisfied give up at once (a test of the fallback).  No reference
 * counterpart. */
int adz_plan_set_spin_limit(adz_plan* p, int polls);
int adz_plan_free(adz_plan* p);
/* Project autotuner for the persistent kernel's tile parameters: a coordinate search
 * (staging, tile order, fields per pass, blocks per CU, loop order, XCD chunking), each candidate
 * timed over `reps` executions of a plan built from `items`; the items' destinations are
 * written.  Returns the fastest tune and its time per execution (ms).  No counterpart in
 * the reference (its slices are fixed at ni*nj points, adz_interp_hlt.F90:102). */
int adz_plan_autotune(const adz_item* items, int nitems, int fdtype, int math, int post, int reps,
                      adz_stream stream, adz_tune* best, float* best_ms);

/* ------------------------------------------------------------------------------------
 * Level 4: Bermejo-Conde mass fixer (Bermejo_Conde.F90:18-320 with the sums of
 * mass_tr_PM.F90 / mass_tr_X.F90) on device arrays.
 * ----------------------------------------------------------------------------------*/

/* One tracer registered for the fixer (type bc, adz_mem.F90:109-125).  DEVICE pointers to
 * element (1,1,1) of model-layout arrays of the fixer's storage type. */
typedef struct adz_bc_tracer {
  const void* p;   /* tracer at time P (tracers_P)                            */
  void* m;         /* tracer at time M after interpolation, corrected in place */
  const void* lin; /* Adz_post lin (already clipped to [min,max])             */
  const void* mn;  /* Adz_post min                                            */
  const void* mx;  /* Adz_post max                                            */
  int legacy;      /* mass = 1 (additive, p = 1)                              */
  int weight;      /* 1 additive, 2 multiplicative, 3 additive * pr_k/pr_s    */
  int pexp_n;      /* index (1-based) into pexp_list when not legacy          */
} adz_bc_tracer;

typedef struct adz_bc_domain {
  int minx, maxx, miny, maxy, nk; /* array layout (l_minx:l_maxx, l_miny:l_maxy, nk)   */
  int i0, in, j0, jn, k0;         /* scope of the sums and corrections (core, k0t:nk)  */
  int min_max;                    /* Adz_BC_min_max_L                                  */
  float pexp_list[9];             /* Adz_BC_pexp_list                                  */
  int autobar;                    /* Schm_autobar_L: every mass is the sum over level 1
                                     of tracer * mask with `mask` = geomh_area_mask_8
                                     (mass_tr_PM.F90:79-85, mass_tr_X.F90:73-79); the
                                     air masses are read but not used; needs k0 = 1   */
} adz_bc_domain;

typedef struct adz_bc adz_bc;
/* Register ntr tracers (tdtype ADZ_F32 = the reference's `real` arrays, or ADZ_F64) and
 * allocate the fixer's scratch (per-block FP64 partial sums, per-tracer scalars) once. */
int adz_bc_create(const adz_bc_tracer* tr, int ntr, const adz_bc_domain* dom, int tdtype, adz_bc** out);
/* Apply the fixer to every registered tracer.  airm1/airm0 (air mass at time P / M) and
 * pkps (pr_k/pr_s, weight 3 only; may be NULL otherwise) are DEVICE arrays of the model
 * layout; mask is the DEVICE FP64 (minx:maxx, miny:maxy) geomh_mask_8; deficit is a DEVICE
 * FP64 array [ntr] holding BC_mass_deficit, read and updated (the carried epsilon,
 * Bermejo_Conde.F90:100,292-308).  diag (DEVICE [6*ntr] or NULL) receives, per tracer:
 * target mass, mass before correction, weight mass, final mass, skipped flag, deficit.
 * Global FP64 sums are computed in a fixed order with double-double partials, so the
 * result is deterministic and equals the exactly rounded sum (the reference sums each
 * rank sequentially, mass_tr_X.F90:86-93).  No host synchronisation: capturable. */
int adz_bermejo_conde(adz_bc* bc, const void* airm1, const void* airm0, const void* pkps, const double* mask,
                      double* deficit, double* diag, adz_stream stream);
/* The same with the LAM boundary flux (Adz_BC_LAM_flux = 1): flux_mass is a DEVICE FP64
 * array [2*ntr] of (mass_fo, mass_fi) per tracer, as adz_bc_lam_flux writes it; the target
 * mass becomes (mass_p - deficit) + (mass_fi - mass_fo) (Bermejo_Conde.F90:92-105). */
int adz_bermejo_conde_lam(adz_bc* bc, const void* airm1, const void* airm0, const void* pkps, const double* mask,
                          const double* flux_mass, double* deficit, double* diag, adz_stream stream);
int adz_bc_free(adz_bc* bc);
/* The fixer in one call that allocates nothing: adz_bc_create + adz_bermejo_conde (flux_mass
 * NULL) or adz_bermejo_conde_lam (flux_mass given) + adz_bc_free, with the same arguments
 * and bit for bit the same results.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under hipGraph capture, tight serving loops).
 * The tracer table, block partials and per-tracer scalars live in a fixed device array of
 * 8.4 MB that each GPU holds until the process ends; at most 64 tracers per call (more
 * return ADZ_EINVAL).  Every call on one GPU uses that GPU's array: calls that overlap on
 * one GPU (on different streams, or in graphs replayed concurrently) race. */
int adz_bermejo_conde_static(const adz_bc_tracer* tr, int ntr, const adz_bc_domain* dom, int tdtype,
                             const void* airm1, const void* airm0, const void* pkps, const double* mask,
                             const double* flux_mass, double* deficit, double* diag, adz_stream stream);

/* ------------------------------------------------------------------------------------
 * Level 4b: LAM boundary fluxes of the fixer, Aranami et al. (2015), Adz_BC_LAM_flux = 1:
 * adz_BC_LAM_Aranami_hlt.F90:18-236 with the masks of adz_BC_LAM_setup.F90:18-92 and the
 * flux masses of mass_tr_PM.F90:101-127.  For a tile that touches the four LAM edges
 * (the one-tile case): mask_o is 1 on the core, mask_i is 1 on 1..ni x 1..nj x 1..nk
 * outside it, both are 0 in the halo (where the reference's work arrays are never
 * written and the clamped positions never reach).
 * ----------------------------------------------------------------------------------*/

/* One fixer tracer: DEVICE pointers to element (1,1,1) of model-layout arrays. */
typedef struct adz_flux_tracer {
  const void* src; /* the tracer interpolated (tracers_P, the source of adz_tricub_hlt) */
  void* fo;        /* Adz_flux%fo (FLUX_out), or NULL: the flux arrays are not stored   */
  void* fi;        /* Adz_flux%fi (FLUX_in), or NULL                                    */
} adz_flux_tracer;

typedef struct adz_flux_domain {
  int minx, maxx, miny, maxy, nk; /* array layout (l_minx:l_maxx, l_miny:l_maxy, nk)            */
  int ni, nj;                     /* the tile l_ni x l_nj                                       */
  int i0b, inb, j0b, jnb;         /* targets Adz_i0b:Adz_inb x Adz_j0b:Adz_jnb x 1:nk (Adz_pb)  */
  int ic0, ic1, jc0, jc1, kc0;    /* core 1+pil_w:l_ni-pil_e x 1+pil_s:l_nj-pil_n x Adz_k0t:nk  */
  int autobar;                    /* Schm_autobar_L: the flux masses are sums over level 1 of
                                     FLUX * mask with `mask` = geomh_area_mask_8
                                     (mass_tr_PM.F90:105-112); the air mass is read, unused     */
} adz_flux_domain;

typedef struct adz_flux adz_flux;
/* Register ntr tracers (tdtype ADZ_F32 or ADZ_F64) and allocate the scratch (a class byte
 * and a list slot per target, per-slice partial sums).  geom is the thermo-level
 * descriptor (Adz_cpntr_t); the tracers are read on their model layout as adz_tricub_hlt
 * reads non-extended sources.  pdtype is the storage type of the Adz_pb positions. */
int adz_flux_create(const adz_flux_tracer* tr, int ntr, const adz_flux_domain* dom, const adz_ztab* geom, int tdtype,
                    int pdtype, adz_flux** out);
/* FLUX_out / FLUX_in of every registered tracer at the positions xyz (DEVICE Adz_pb), in
 * the reference's arithmetic (ADZ_MATH_REF), stored when the tracers carry fo/fi arrays,
 * and their masses with the air mass at time M (airm0, DEVICE, element (1,1,1)) and the
 * DEVICE geomh_mask_8 (element (1,1)) into mass (DEVICE FP64 [2*ntr]: mass_fo, mass_fi),
 * summed as adz_bermejo_conde sums.  Only the targets whose stencil crosses the core's
 * boundary are interpolated; the others hold zeros.  Capturable (no host synchronisation). */
int adz_bc_lam_flux(adz_flux* fx, const void* xyz, const void* airm0, const double* mask, double* mass,
                    adz_stream stream);
/* The ZLF form (Adz_BC_LAM_flux = 2): the element-wise steps of adz_BC_LAM_zlf_hlt
 * (adz_BC_LAM_zlf_hlt.F90:18-319) on nf DEVICE arrays (HOST tables f1 / f2 of pointers to
 * element (1,1,1) of model-layout arrays) over the tile of a one-tile LAM:
 *   1: f1 (a source) = 0 outside the extension (pilot zones minus ext = Grd_maxcfl + 1
 *      cells) and, inside it, below level nint(k0/2);
 *   2: f2 (the piloting copy, tracers_B) = f1 (the destination) outside the core;
 *   3: f1 = 0 outside [5, ni-4] x [5, nj-4];
 *   4: f1 = f2 outside the core.
 * The core is 1+pil_w:ni-pil_e x 1+pil_s:nj-pil_n x k0:nk with k0 = Adz_k0t.  The caller
 * runs them as adz_BC_LAM_zlf_0_hlt does (adz_BC_LAM_zlf_0_hlt.F90:55-95): 1 and 2 before
 * interpolating on Adz_pb, 3 before the fixer (whose scope is then Adz_i0b:Adz_inb x
 * Adz_j0b:Adz_jnb x 1:nk, adz_post_tr.F90:147-151), 4 after it. */
typedef struct adz_zlf_domain {
  int minx, maxx, miny, maxy, nk; /* array layout                      */
  int ni, nj;                     /* the tile                          */
  int pil_w, pil_e, pil_s, pil_n; /* Glb_pil_* of the four LAM edges   */
  int ext;                        /* Grd_maxcfl + 1                    */
  int k0;                         /* Adz_k0t                           */
} adz_zlf_domain;
int adz_bc_lam_zlf(int setup, void* const* f1, void* const* f2, int nf, const adz_zlf_domain* dom, int dtype,
                   adz_stream stream);
/* Number of targets interpolated by the last adz_bc_lam_flux (synchronises; diagnostics).
 * No counterpart in the reference, which interpolates every target. */
int adz_flux_active(const adz_flux* fx, int* count);
int adz_flux_free(adz_flux* fx);

/* ------------------------------------------------------------------------------------
 * Level 4c: ILMC shape preservation (mono = 2): ILMC_LAM (ILMC_LAM.F90:18-273) with its
 * sweep loop (ilmc_lam_loop.inc:16-139), called from adz_post_tr.F90:104 for each tracer
 * whose mono code is 2.  For a tile that touches the four LAM edges: the domain is
 * i0..in x j0..jn x k0..nk (Adz_i0..Adz_in x Adz_j0..Adz_jn x Adz_k0t..nk,
 * ILMC_LAM.F90:81-86).  Arrays are FP32 (`real`) in the model layout; pointers address
 * element (1,1,1). */
typedef struct adz_ilmc_tracer {
  void* f;        /* Adz_stack(n)%dst after interpolation, corrected in place */
  const void* mn; /* Adz_post(n)%min                                          */
  const void* mx; /* Adz_post(n)%max                                          */
} adz_ilmc_tracer;

typedef struct adz_ilmc_domain {
  int minx, maxx, miny, maxy, nk; /* array layout (l_minx:l_maxx, l_miny:l_maxy, nk) */
  int i0, in, j0, jn, k0;         /* the domain                                     */
  int sweep_max;                  /* Adz_ILMC_sweep_max, 1..4 (reference default 2) */
  int min_max;                    /* Adz_ILMC_min_max_L (reference default true)    */
  int autobar;                    /* Schm_autobar_L: sweeps stay on one level       */
} adz_ilmc_domain;

typedef struct adz_ilmc adz_ilmc;
/* Register ntr tracers and allocate the visiting lists once (8 bytes per cell of the
 * domain per tracer, the worst case of every cell out of bounds).  Replaces the per-cell
 * sweep lists the reference allocates on its first call (ILMC_LAM.F90:138-220); the sweep
 * cells are enumerated in the kernel instead. */
int adz_ilmc_create(const adz_ilmc_tracer* tr, int ntr, const adz_ilmc_domain* dom, adz_ilmc** out);
/* Apply ILMC to every registered tracer; air_m is the DEVICE FP32 air mass at time M
 * (airm0).  The result equals the reference's sequential (k, j, i) visit bit for bit.
 * counts (DEVICE int64 [3*ntr] or NULL) receives per tracer: sweeps that moved mass
 * (reset(:,5)), min resets and max resets (reset(:,3:4)).  Tracer arrays must not
 * overlap, and calls on one plan must not overlap (they share its visiting lists).  No host
 * synchronisation: capturable (one launch per hyperplane of the domain, about
 * ni + (2S+1) nj + (2S+1)^2 nk launches). */
int adz_ilmc_lam(adz_ilmc* il, const void* air_m, long long* counts, adz_stream stream);
int adz_ilmc_free(adz_ilmc* il);

/* ------------------------------------------------------------------------------------
 * Level 4d: the fixer's air masses.  get_air_mass_hlt (get_air_mass_hlt.F90:18-103) at
 * time P (F_time = 1) and time M (F_time = 0), and pr_t/pr_s at time M, as set_post_tr_hlt
 * forms them before the tracers' post-processing (set_post_tr_hlt.F90:44-70).
 * ----------------------------------------------------------------------------------*/
typedef struct adz_airmass_fields {
  int minx, maxx, miny, maxy;  /* layout of every array: (l_minx:l_maxx, l_miny:l_maxy, .)   */
  const void* pm_p;            /* pw_pm_plus  (nk+1 levels): momentum-level pressure at P     */
  const void* pm_m;            /* pw_pm_moins (nk+1 levels), at M                            */
  const void* hu_p;            /* tracers_P(Tr3d_hu): read when dry_mixing_ratio             */
  const void* hu_m;            /* tracers_M(Tr3d_hu)                                          */
  const void* const* hyd_p;    /* HOST array of nhyd DEVICE pointers: the Tr3d_wload tracers
                                  of trt1 in tracer order; read when dry_mixing_ratio and
                                  wload (sumhydro_hlt.F90:45-60)                              */
  const void* const* hyd_m;    /* the same of trt0                                            */
  int nhyd;                    /* 0..8                                                        */
  const void* pt_m;            /* pw_pt_moins (nk levels)                                     */
  const void* p0_m;            /* pw_p0_moins (2-D)                                           */
  const double* area;          /* geomh_area_8 (2-D, FP64)                                    */
  int dry_mixing_ratio;        /* Schm_dry_mixing_ratio_L                                     */
  int wload;                   /* Schm_wload_L                                                */
} adz_airmass_fields;
/* airm1 (time P), airm0 (time M) and pkps (pr_t/pr_s at M, may be NULL) of storage type
 * dtype (ADZ_F32 = the reference's `real`, or ADZ_F64; every input array has that type
 * except area), DEVICE pointers to element (1,1,1) of the layout.  The air masses are
 * written on the whole layout (zero outside 1..l_ni x 1..l_nj and below Adz_k0t, as the
 * reference does); pkps on 1..l_ni x 1..l_nj x 1..nk.  The layer thicknesses
 * Ver_dz_8%t / Ver_idz_8%t come from the context's levels (set_zeta.F90:271-275);
 * grav is tdpack's grav_8.  No host synchronisation: capturable. */
int adz_set_post_tr(const adz_ctx* c, const adz_airmass_fields* f, double grav, void* airm1, void* airm0, void* pkps,
                    int dtype, adz_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* ADZ_HIP_H */
