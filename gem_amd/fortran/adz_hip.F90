! adz_hip.F90 -- ISO_C_BINDING interface to libadz_hip.so (declared in include/adz_hip.h).
!
! One bind(C) interface per entry point and one bind(C) derived type per struct, with the
! C names unchanged, so Fortran host code calls the HIP kernels with no wrapper code in
! between.  The reference's own binding of its C kernels is the same kind of file
! (src/gemdyn/src/base/tricublin_f90.inc:1-80, bind(C) interfaces to turbo_tricublin.c);
! the module adz_gpu (adz_gpu.F90) builds the reference's adz_* signatures on top of it.
!
! Device arrays are passed as type(c_ptr) holding HIP device addresses (hipMalloc'd, or
! the device address of an OpenMP-offloaded array).  Host arrays are passed by reference.
module adz_hip
   use, intrinsic :: iso_c_binding
   implicit none
   public

   integer(c_int), parameter :: ADZ_OK = 0, ADZ_EINVAL = 1, ADZ_EHIP = 2, ADZ_ENOMEM = 3, ADZ_ESTATE = 4
   integer(c_int), parameter :: ADZ_F64 = 0, ADZ_F32 = 1, ADZ_BF16 = 2
   integer(c_int), parameter :: ADZ_MATH_REF = 0, ADZ_MATH_F32 = 1
   integer(c_int), parameter :: ADZ_POST_STORE = 1, ADZ_POST_CLIP = 2
   integer(c_int), parameter :: ADZ_A_PXYZM = 0, ADZ_A_WPXYZ = 1, ADZ_A_UVW_DEP = 2, ADZ_A_PM = 3, &
                                ADZ_A_PMU = 4, ADZ_A_PMV = 5, ADZ_A_PT = 6, ADZ_A_PB = 7, ADZ_A_UVW_D = 8, &
                                ADZ_A_WARR = 9

   type, bind(C) :: adz_grid
      integer(c_int) :: ni, nj, nk, i0g, j0g, halox, haloy
      integer(c_int) :: west, east, south, north, pil_w, pil_e, pil_s, pil_n
      integer(c_int) :: gbpil_t, bc_lam_flux, niu, njv
      real(c_double) :: inv_hx, inv_hy, inv_rayt
      type(c_ptr)    :: zm, zt, zx, lat        ! HOST arrays (c_loc)
      real(c_double) :: zmin, zmax, wmstar_nk, idz_t_nkm1
   end type adz_grid

   type, bind(C) :: adz_field
      type(c_ptr) :: src = c_null_ptr, dst = c_null_ptr, lin = c_null_ptr, mn = c_null_ptr, mx = c_null_ptr
   end type adz_field

   type, bind(C) :: adz_item
      type(c_ptr)    :: geom = c_null_ptr, xyz = c_null_ptr
      integer(c_int) :: pdtype = ADZ_F32
      integer(c_int) :: i0 = 0, in = 0, j0 = 0, jn = 0, k0 = 0, nk = 0
      integer(c_int) :: dminx = 0, dmaxx = 0, dminy = 0, dmaxy = 0, ext = 0
      type(c_ptr)    :: fld = c_null_ptr       ! HOST array of adz_field (c_loc)
      integer(c_int) :: nptr = 0
   end type adz_item

   type, bind(C) :: adz_tune
      integer(c_int) :: fields_per_pass = 0, points_outer = 0, blocks_per_cu = 0, xcd_chunks = 1, stage = 0
      integer(c_int) :: level_inner = 0
   end type adz_tune

   type, bind(C) :: adz_bc_tracer
      type(c_ptr)    :: p = c_null_ptr, m = c_null_ptr, lin = c_null_ptr, mn = c_null_ptr, mx = c_null_ptr
      integer(c_int) :: legacy = 1, weight = 1, pexp_n = 0
   end type adz_bc_tracer

   type, bind(C) :: adz_bc_domain
      integer(c_int) :: minx, maxx, miny, maxy, nk
      integer(c_int) :: i0, in, j0, jn, k0
      integer(c_int) :: min_max
      real(c_float)  :: pexp_list(9)
      integer(c_int) :: autobar = 0      ! Schm_autobar_L: level-1 masses, mask = geomh_area_mask_8
   end type adz_bc_domain

   type, bind(C) :: adz_flux_tracer
      type(c_ptr) :: src = c_null_ptr, fo = c_null_ptr, fi = c_null_ptr
   end type adz_flux_tracer

   type, bind(C) :: adz_zlf_domain
      integer(c_int) :: minx, maxx, miny, maxy, nk
      integer(c_int) :: ni, nj
      integer(c_int) :: pil_w, pil_e, pil_s, pil_n
      integer(c_int) :: ext, k0
   end type adz_zlf_domain

   type, bind(C) :: adz_ilmc_tracer
      type(c_ptr) :: f = c_null_ptr, mn = c_null_ptr, mx = c_null_ptr
   end type adz_ilmc_tracer

   type, bind(C) :: adz_ilmc_domain
      integer(c_int) :: minx, maxx, miny, maxy, nk
      integer(c_int) :: i0, in, j0, jn, k0
      integer(c_int) :: sweep_max, min_max, autobar
   end type adz_ilmc_domain

   type, bind(C) :: adz_flux_domain
      integer(c_int) :: minx, maxx, miny, maxy, nk
      integer(c_int) :: ni, nj
      integer(c_int) :: i0b, inb, j0b, jnb
      integer(c_int) :: ic0, ic1, jc0, jc1, kc0
      integer(c_int) :: autobar = 0      ! Schm_autobar_L
   end type adz_flux_domain

   type, bind(C) :: adz_airmass_fields
      integer(c_int) :: minx = 0, maxx = 0, miny = 0, maxy = 0
      type(c_ptr)    :: pm_p = c_null_ptr, pm_m = c_null_ptr, hu_p = c_null_ptr, hu_m = c_null_ptr
      type(c_ptr)    :: hyd_p = c_null_ptr, hyd_m = c_null_ptr  ! HOST arrays of nhyd device addresses
      integer(c_int) :: nhyd = 0
      type(c_ptr)    :: pt_m = c_null_ptr, p0_m = c_null_ptr, area = c_null_ptr
      integer(c_int) :: dry_mixing_ratio = 0, wload = 0
   end type adz_airmass_fields

   interface
      function adz_last_error() bind(C, name="adz_last_error") result(msg)
         import :: c_ptr
         type(c_ptr) :: msg
      end function
      function adz_version() bind(C, name="adz_version") result(v)
         import :: c_int
         integer(c_int) :: v
      end function

      ! ---------------------------------------------------------------- level 1
      function adz_vsearch_setup_plus(levels, nk, ni, nj, offseti, offsetj, lv) &
            bind(C, name="adz_vsearch_setup_plus") result(rc)
         import :: c_int, c_double, c_ptr
         real(c_double), intent(in) :: levels(*)
         integer(c_int), value :: nk, ni, nj, offseti, offsetj
         type(c_ptr), intent(out) :: lv
         integer(c_int) :: rc
      end function
      function adz_vsearch_set_bounds(lv, ilo, ihi, jlo, jhi) bind(C, name="adz_vsearch_set_bounds") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: lv
         integer(c_int), value :: ilo, ihi, jlo, jhi
         integer(c_int) :: rc
      end function
      function adz_vsearch_free(lv) bind(C, name="adz_vsearch_free") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: lv
         integer(c_int) :: rc
      end function
      function adz_vcoef(ijk, zlin, coef, pxyz, pdtype, lv, n, stream) bind(C, name="adz_vcoef") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ijk, zlin, coef, pxyz, lv, stream
         integer(c_int), value :: pdtype, n
         integer(c_int) :: rc
      end function
      function adz_tricublin_zyx1_m_n(d, fs, pxyz, lv, n, m, fdtype, pdtype, math, stream) &
            bind(C, name="adz_tricublin_zyx1_m_n") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: d, pxyz, lv, stream
         type(c_ptr), intent(in) :: fs(*)
         integer(c_int), value :: n, m, fdtype, pdtype, math
         integer(c_int) :: rc
      end function
      function adz_tricublin_mono_zyx_m_n(d, l, mi, ma, fs, pxyz, lv, n, m, fdtype, pdtype, math, stream) &
            bind(C, name="adz_tricublin_mono_zyx_m_n") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: d, l, mi, ma, pxyz, lv, stream
         type(c_ptr), intent(in) :: fs(*)
         integer(c_int), value :: n, m, fdtype, pdtype, math
         integer(c_int) :: rc
      end function
      function adz_tricublin_zyx3_n(d, f3, pxyz, lv, n, fdtype, pdtype, math, stream) &
            bind(C, name="adz_tricublin_zyx3_n") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: d, f3, pxyz, lv, stream
         integer(c_int), value :: n, fdtype, pdtype, math
         integer(c_int) :: rc
      end function

      ! ---------------------------------------------------------------- level 2
      function adz_set(g, pos_dtype, itraj, ctx) bind(C, name="adz_set") result(rc)
         import :: c_int, c_ptr, adz_grid
         type(adz_grid), intent(in) :: g
         integer(c_int), value :: pos_dtype, itraj
         type(c_ptr), intent(out) :: ctx
         integer(c_int) :: rc
      end function
      function adz_ctx_free(ctx) bind(C, name="adz_ctx_free") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int) :: rc
      end function
      function adz_ctx_info(ctx, ints, dbls) bind(C, name="adz_ctx_info") result(rc)
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: ints(24)
         real(c_double), intent(out) :: dbls(9)
         integer(c_int) :: rc
      end function
      function adz_ctx_table(ctx, which, host_out) bind(C, name="adz_ctx_table") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, host_out
         integer(c_int), value :: which
         integer(c_int) :: rc
      end function
      function adz_ctx_array(ctx, which, dptr, nelem) bind(C, name="adz_ctx_array") result(rc)
         import :: c_int, c_long, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: which
         type(c_ptr), intent(out) :: dptr
         integer(c_long), intent(out) :: nelem
         integer(c_int) :: rc
      end function
      function adz_ctx_ztab(ctx, thermo) bind(C, name="adz_ctx_ztab") result(lv)
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: thermo
         type(c_ptr) :: lv
      end function
      ! restart of the SL state (Adz_pxyzm, Adz_wpxyz, next iteration count) via a host buffer
      function adz_ctx_state_bytes(ctx, bytes) bind(C, name="adz_ctx_state_bytes") result(rc)
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: ctx
         integer(c_size_t), intent(out) :: bytes
         integer(c_int) :: rc
      end function
      function adz_ctx_save_state(ctx, host, bytes, stream) bind(C, name="adz_ctx_save_state") result(rc)
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: ctx, host, stream
         integer(c_size_t), value :: bytes
         integer(c_int) :: rc
      end function
      function adz_ctx_load_state(ctx, host, bytes, stream) bind(C, name="adz_ctx_load_state") result(rc)
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: ctx, host, stream
         integer(c_size_t), value :: bytes
         integer(c_int) :: rc
      end function
      function adz_inittraj(ctx, stream) bind(C, name="adz_inittraj") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx, stream
         integer(c_int) :: rc
      end function
      function adz_traject(ctx, dt, schm_advec, bA_m, bA, uu_ext, vv_ext, ww_ext, ut0, vt0, zdt0, stream) &
            bind(C, name="adz_traject") result(rc)
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx, uu_ext, vv_ext, ww_ext, ut0, vt0, zdt0, stream
         real(c_double), value :: dt, bA_m, bA
         integer(c_int), value :: schm_advec
         integer(c_int) :: rc
      end function

      ! ---------------------------------------------------------------- level 3
      function adz_tricub_hlt(fld, nptr, xyz, pdtype, geom, i0, in, j0, jn, k0, nk, dminx, dmaxx, dminy, dmaxy, &
                              ext, post, fdtype, math, stream) bind(C, name="adz_tricub_hlt") result(rc)
         import :: c_int, c_ptr, adz_field
         type(adz_field), intent(in) :: fld(*)
         type(c_ptr), value :: xyz, geom, stream
         integer(c_int), value :: nptr, pdtype, i0, in, j0, jn, k0, nk, dminx, dmaxx, dminy, dmaxy, ext, post, &
                                  fdtype, math
         integer(c_int) :: rc
      end function
      function adz_tricub_hlt_qv(ctx, fld, nptr, xyz, i0, in, j0, jn, k0, dminx, dmaxx, dminy, dmaxy, ext, post, &
                                 stream) bind(C, name="adz_tricub_hlt_qv") result(rc)
         import :: c_int, c_ptr, adz_field
         type(c_ptr), value :: ctx, xyz, stream
         type(adz_field), intent(in) :: fld(*)
         integer(c_int), value :: nptr, i0, in, j0, jn, k0, dminx, dmaxx, dminy, dmaxy, ext, post
         integer(c_int) :: rc
      end function
      function adz_plan_create(items, nitems, fdtype, math, post, tune, plan) bind(C, name="adz_plan_create") &
            result(rc)
         import :: c_int, c_ptr, adz_item
         type(adz_item), intent(in) :: items(*)
         integer(c_int), value :: nitems, fdtype, math, post
         type(c_ptr), value :: tune               ! c_loc(adz_tune) or c_null_ptr
         type(c_ptr), intent(out) :: plan
         integer(c_int) :: rc
      end function
      function adz_plan_execute(plan, stream) bind(C, name="adz_plan_execute") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: plan, stream
         integer(c_int) :: rc
      end function
      function adz_plan_tune(plan, tune) bind(C, name="adz_plan_tune") result(rc)
         import :: c_int, c_ptr, adz_tune
         type(c_ptr), value :: plan
         type(adz_tune), intent(out) :: tune
         integer(c_int) :: rc
      end function
      function adz_plan_free(plan) bind(C, name="adz_plan_free") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: plan
         integer(c_int) :: rc
      end function
      function adz_plan_autotune(items, nitems, fdtype, math, post, reps, stream, best, best_ms) &
            bind(C, name="adz_plan_autotune") result(rc)
         import :: c_int, c_float, c_ptr, adz_item, adz_tune
         type(adz_item), intent(in) :: items(*)
         integer(c_int), value :: nitems, fdtype, math, post, reps
         type(c_ptr), value :: stream
         type(adz_tune), intent(out) :: best
         real(c_float), intent(out) :: best_ms
         integer(c_int) :: rc
      end function

      ! ---------------------------------------------------------------- level 4
      function adz_bc_create(tr, ntr, dom, tdtype, bc) bind(C, name="adz_bc_create") result(rc)
         import :: c_int, c_ptr, adz_bc_tracer, adz_bc_domain
         type(adz_bc_tracer), intent(in) :: tr(*)
         integer(c_int), value :: ntr, tdtype
         type(adz_bc_domain), intent(in) :: dom
         type(c_ptr), intent(out) :: bc
         integer(c_int) :: rc
      end function
      function adz_bermejo_conde(bc, airm1, airm0, pkps, mask, deficit, diag, stream) &
            bind(C, name="adz_bermejo_conde") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: bc, airm1, airm0, pkps, mask, deficit, diag, stream
         integer(c_int) :: rc
      end function
      function adz_bermejo_conde_lam(bc, airm1, airm0, pkps, mask, flux_mass, deficit, diag, stream) &
            bind(C, name="adz_bermejo_conde_lam") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: bc, airm1, airm0, pkps, mask, flux_mass, deficit, diag, stream
         integer(c_int) :: rc
      end function
      function adz_bc_free(bc) bind(C, name="adz_bc_free") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: bc
         integer(c_int) :: rc
      end function

      ! ---------------------------------------------------------------- level 4b
      function adz_flux_create(tr, ntr, dom, geom, tdtype, pdtype, fx) bind(C, name="adz_flux_create") result(rc)
         import :: c_int, c_ptr, adz_flux_tracer, adz_flux_domain
         type(adz_flux_tracer), intent(in) :: tr(*)
         integer(c_int), value :: ntr, tdtype, pdtype
         type(adz_flux_domain), intent(in) :: dom
         type(c_ptr), value :: geom
         type(c_ptr), intent(out) :: fx
         integer(c_int) :: rc
      end function
      function adz_bc_lam_flux(fx, xyz, airm0, mask, mass, stream) bind(C, name="adz_bc_lam_flux") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: fx, xyz, airm0, mask, mass, stream
         integer(c_int) :: rc
      end function
      function adz_flux_active(fx, count) bind(C, name="adz_flux_active") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: fx
         integer(c_int), intent(out) :: count
         integer(c_int) :: rc
      end function
      function adz_bc_lam_zlf(setup, f1, f2, nf, dom, dtype, stream) bind(C, name="adz_bc_lam_zlf") result(rc)
         import :: c_int, c_ptr, adz_zlf_domain
         integer(c_int), value :: setup, nf, dtype
         type(c_ptr), intent(in) :: f1(*), f2(*)
         type(adz_zlf_domain), intent(in) :: dom
         type(c_ptr), value :: stream
         integer(c_int) :: rc
      end function
      function adz_ilmc_create(tr, ntr, dom, il) bind(C, name="adz_ilmc_create") result(rc)
         import :: c_int, c_ptr, adz_ilmc_tracer, adz_ilmc_domain
         type(adz_ilmc_tracer), intent(in) :: tr(*)
         integer(c_int), value :: ntr
         type(adz_ilmc_domain), intent(in) :: dom
         type(c_ptr), intent(out) :: il
         integer(c_int) :: rc
      end function
      function adz_ilmc_lam(il, air_m, counts, stream) bind(C, name="adz_ilmc_lam") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: il, air_m, counts, stream
         integer(c_int) :: rc
      end function
      function adz_ilmc_free(il) bind(C, name="adz_ilmc_free") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: il
         integer(c_int) :: rc
      end function
      function adz_flux_free(fx) bind(C, name="adz_flux_free") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: fx
         integer(c_int) :: rc
      end function
      function adz_set_post_tr(ctx, f, grav, airm1, airm0, pkps, dtype, stream) bind(C, name="adz_set_post_tr") &
         result(rc)
         import :: c_int, c_ptr, c_double, adz_airmass_fields
         type(c_ptr), value :: ctx
         type(adz_airmass_fields), intent(in) :: f
         real(c_double), value :: grav
         type(c_ptr), value :: airm1, airm0, pkps
         integer(c_int), value :: dtype
         type(c_ptr), value :: stream
         integer(c_int) :: rc
      end function
      function adz_plan_status(plan, aborted) bind(C, name="adz_plan_status") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: plan
         integer(c_int), intent(out) :: aborted
         integer(c_int) :: rc
      end function
   end interface

   ! The few HIP runtime calls the Fortran layer needs (libamdhip64).
   integer(c_int), parameter :: hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2
   interface
      function hipMalloc(ptr, bytes) bind(C, name="hipMalloc") result(rc)
         import :: c_int, c_size_t, c_ptr
         type(c_ptr), intent(out) :: ptr
         integer(c_size_t), value :: bytes
         integer(c_int) :: rc
      end function
      function hipFree(ptr) bind(C, name="hipFree") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ptr
         integer(c_int) :: rc
      end function
      function hipMemcpy(dst, src, bytes, kind) bind(C, name="hipMemcpy") result(rc)
         import :: c_int, c_size_t, c_ptr
         type(c_ptr), value :: dst, src
         integer(c_size_t), value :: bytes
         integer(c_int), value :: kind
         integer(c_int) :: rc
      end function
      function hipDeviceSynchronize() bind(C, name="hipDeviceSynchronize") result(rc)
         import :: c_int
         integer(c_int) :: rc
      end function
      function hipMemcpyAsync(dst, src, bytes, kind, stream) bind(C, name="hipMemcpyAsync") result(rc)
         import :: c_int, c_size_t, c_ptr
         type(c_ptr), value :: dst, src
         integer(c_size_t), value :: bytes
         integer(c_int), value :: kind
         type(c_ptr), value :: stream
         integer(c_int) :: rc
      end function
      function hipStreamSynchronize(stream) bind(C, name="hipStreamSynchronize") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: stream
         integer(c_int) :: rc
      end function
      function hipStreamCreate(stream) bind(C, name="hipStreamCreate") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), intent(out) :: stream
         integer(c_int) :: rc
      end function
      function hipStreamDestroy(stream) bind(C, name="hipStreamDestroy") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: stream
         integer(c_int) :: rc
      end function
      ! graph capture of the work enqueued on a stream (mode 0: global, the strictest)
      function hipStreamBeginCapture(stream, mode) bind(C, name="hipStreamBeginCapture") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: stream
         integer(c_int), value :: mode
         integer(c_int) :: rc
      end function
      function hipStreamEndCapture(stream, graph) bind(C, name="hipStreamEndCapture") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: stream
         type(c_ptr), intent(out) :: graph
         integer(c_int) :: rc
      end function
      function hipGraphInstantiate(exec, graph, errnode, logbuf, bufsize) bind(C, name="hipGraphInstantiate") &
         result(rc)
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), intent(out) :: exec
         type(c_ptr), value :: graph, errnode, logbuf
         integer(c_size_t), value :: bufsize
         integer(c_int) :: rc
      end function
      function hipGraphLaunch(exec, stream) bind(C, name="hipGraphLaunch") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: exec, stream
         integer(c_int) :: rc
      end function
      function hipGraphExecDestroy(exec) bind(C, name="hipGraphExecDestroy") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: exec
         integer(c_int) :: rc
      end function
      function hipGraphDestroy(graph) bind(C, name="hipGraphDestroy") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: graph
         integer(c_int) :: rc
      end function
      function hipEventCreate(ev) bind(C, name="hipEventCreate") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), intent(out) :: ev
         integer(c_int) :: rc
      end function
      function hipEventRecord(ev, stream) bind(C, name="hipEventRecord") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ev, stream
         integer(c_int) :: rc
      end function
      function hipEventSynchronize(ev) bind(C, name="hipEventSynchronize") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ev
         integer(c_int) :: rc
      end function
      function hipEventElapsedTime(ms, ev0, ev1) bind(C, name="hipEventElapsedTime") result(rc)
         import :: c_int, c_ptr, c_float
         real(c_float), intent(out) :: ms
         type(c_ptr), value :: ev0, ev1
         integer(c_int) :: rc
      end function
      function hipEventDestroy(ev) bind(C, name="hipEventDestroy") result(rc)
         import :: c_int, c_ptr
         type(c_ptr), value :: ev
         integer(c_int) :: rc
      end function
   end interface

contains

   ! The message of the last failed adz_* call, as a Fortran string.
   function adz_error_message() result(msg)
      character(len=:), allocatable :: msg
      character(kind=c_char), pointer :: s(:)
      type(c_ptr) :: p
      integer :: n
      p = adz_last_error()
      if (.not. c_associated(p)) then
         msg = ''
         return
      end if
      call c_f_pointer(p, s, [4096])
      n = 0
      do while (n < 4096)
         if (s(n + 1) == c_null_char) exit
         n = n + 1
      end do
      allocate (character(len=n) :: msg)
      msg = transfer(s(1:n), msg)
   end function adz_error_message

end module adz_hip
