! adz_gpu.F90 -- the reference's adz_* entry points with their Fortran signatures, as thin
! ISO_C_BINDING calls into libadz_hip.so on device-resident fields and positions.
!
! Drop-in map (reference file:line -> routine here):
!   adz_set.F90:16-311          -> adz_gpu_set (grid description in, module state out)
!   adz_inittraj.F90:16-42      -> adz_inittraj
!   adz_traject.F90:16-130      -> adz_traject(F_dt_8)         (same signature)
!   adz_main.F90:16-71          -> adz_main()                  (same signature)
!   adz_tracers.F90:16-184      -> adz_tracers(F_before_psadj_L) (same signature)
!   adz_interp_hlt.F90:28-187   -> adz_tricub_hlt(F_stk, F_nptr, F_xyz, F_geom, F_num,
!                                    F_i0, F_in, F_j0, F_jn, F_k0, F_ext_L, F_post, F_QV_L)
!   set_post_tr_hlt.F90:18-70   -> set_post_tr_hlt()           (get_air_mass_hlt at P and M, pkps)
!   adz_post_tr.F90:18-276      -> adz_post_tr(F_wp)            (same signature, F_wp = 0, 1, 2)
!   Bermejo_Conde.F90:18-320    -> Bermejo_Conde(F_bc, F_ntr_bc, F_i0, F_in, F_j0, F_jn, F_k0)
!   adz_BC_LAM_Aranami_hlt.F90:18-236 -> adz_BC_LAM_Aranami_hlt(F_stk, F_post, F_nptr), called by
!                                    adz_tricub_hlt with F_post on a LAM with Adz_BC_LAM_flux = 1
!   adz_BC_LAM_zlf_hlt.F90:18-319, adz_BC_LAM_zlf_0_hlt.F90:18-102 -> adz_BC_LAM_zlf_hlt(F1, F2, F_n,
!                                    F_setup), adz_BC_LAM_zlf_0_hlt(F_nptr, F_number)
!   ILMC_LAM.F90:18-273         -> ILMC_LAM(F_ns, F_name_S, F_i0, F_in, F_j0, F_jn)
!
! The module variables stand in for what the reference keeps in adz_mem / adz_options /
! dyn_fisl_options / cstv / tr3d / mem_tracers / mem_tstp / gmm_pw (same names).
! Array-valued members that the reference declares as real pointers (tracers_P(n)%pntr,
! Adz_pntr_stack%src/dst, post%lin/min/max, bc%p/m/..., rhsu, orhsu_ext, pw_pm_plus, ...)
! are type(c_ptr) device addresses of element (1,1,1) here, and the position sets Adz_pm,
! Adz_pmu, Adz_pmv, Adz_pt, Adz_pb are the device arrays the context owns.
!
! The routines of a timestep (adz_traject, adz_main, adz_tracers, adz_post_tr,
! Bermejo_Conde, set_post_tr_hlt) enqueue work on Adz_stream and never synchronise the
! host, so a caller may capture a whole step in a hipGraph (hipStreamBeginCapture on a
! created stream).  Each batched launch, fixer, boundary-flux and ILMC plan they use is
! created on first use and kept in a cache keyed by its arrays and bounds; a step that is
! captured must therefore have run once before the capture with the same arrays (the
! usual warm-up step), so that the capture meets no allocation.
!
! Differences a caller sees:
!   * F_xyz is the device address of the positions (type(c_ptr)), not a host real array.
!   * With F_post present, tracers whose meta has mono == 1 are clipped to [min, max]
!     inside the interpolation launch (the same operation adz_post_tr.F90:81-94 applies to
!     the same box afterwards), so adz_post_tr(1|2) runs ILMC and registers the fixer's
!     tracers only.  ILMC runs once for all mono = 2 tracers of a group (one launch per
!     hyperplane for all of them); each tracer's result is the sequential loop's, as the
!     tracers do not interact.
!   * The carried mass deficits stay on the device from one Bermejo_Conde call to the
!     next (Adz_bc_deficit).  They are read from F_bc(:)%mass_deficit on the first call and
!     whenever Adz_bc_deficit_upload_L is set (after a restart), and written back into
!     Tr_3CWP / Tr_BQWP(:)%BC_mass_deficit and Adz_bc(:)%mass_deficit by
!     adz_bc_sync_deficits (for output or a restart file).
!   * The LAM boundary flux keeps only what Bermejo_Conde reads of it: the masses of
!     FLUX_out / FLUX_in per tracer (mass_tr_PM.F90:101-127, with the air mass at time M
!     and geomh_mask_8 of the module state when adz_tricub_hlt runs), in device memory
!     (Adz_flux_3CWP / Adz_flux_BQWP, the fixer's tracers of each group in order), not the
!     Adz_flux arrays themselves.
module adz_gpu
   use, intrinsic :: iso_c_binding
   use, intrinsic :: iso_fortran_env, only: REAL64, error_unit
   use adz_hip, c_adz_set => adz_set, c_adz_inittraj => adz_inittraj, c_adz_traject => adz_traject, &
                c_adz_tricub_hlt => adz_tricub_hlt, c_adz_bc_lam_flux => adz_bc_lam_flux
   implicit none
   private :: c_adz_set, c_adz_inittraj, c_adz_traject, c_adz_tricub_hlt, c_adz_bc_lam_flux

   ! ------------------------------------------------------------ types (adz_mem.F90:66-125, tr3d_mod.F90:63-70)
   type :: Adz_pntr_stack
      type(c_ptr) :: src = c_null_ptr, dst = c_null_ptr, pil = c_null_ptr
   end type Adz_pntr_stack

   type :: post
      type(c_ptr) :: lin = c_null_ptr, min = c_null_ptr, max = c_null_ptr
   end type post

   type :: meta_tracers
      character(len=4)  :: name = ''
      character(len=12) :: intp = 'TRICUB'
      logical :: wload = .false., hzd = .false.
      integer :: mono = 0, mass = 0
      real    :: vmin = -huge(1.0), vmax = huge(1.0)
      real(kind=REAL64) :: BC_mass_deficit = 0.d0
   end type meta_tracers

   type :: bc
      integer :: n = 0, wp = 0, pexp_n = 0, weight = 1
      character(len=8) :: name_S = ''
      real(kind=REAL64) :: mass_deficit = 0.d0
      logical :: LEGACY_L = .true.
      type(c_ptr) :: p = c_null_ptr, m = c_null_ptr, w = c_null_ptr, fo = c_null_ptr, fi = c_null_ptr
      type(c_ptr) :: lin = c_null_ptr, min = c_null_ptr, max = c_null_ptr
   end type bc

   ! ------------------------------------------------------------ state (adz_mem / adz_options / ...)
   type(c_ptr) :: Adz_ctx = c_null_ptr
   type(c_ptr) :: Adz_cpntr_q = c_null_ptr, Adz_cpntr_t = c_null_ptr
   type(c_ptr) :: Adz_pm = c_null_ptr, Adz_pmu = c_null_ptr, Adz_pmv = c_null_ptr, Adz_pt = c_null_ptr, &
                  Adz_pb = c_null_ptr, Adz_pxyzm = c_null_ptr, Adz_wpxyz = c_null_ptr
   integer :: Adz_lminx, Adz_lmaxx, Adz_lminy, Adz_lmaxy, Adz_nit, Adz_njt, Adz_nij
   integer :: Adz_i0, Adz_in, Adz_j0, Adz_jn, Adz_i0u, Adz_inu, Adz_j0v, Adz_jnv
   integer :: Adz_i0b, Adz_inb, Adz_j0b, Adz_jnb, Adz_k0, Adz_k0t, Adz_k0m
   integer :: Adz_num_u, Adz_num_v, Adz_num_q, Adz_num_t, Adz_num_b
   integer :: l_minx, l_maxx, l_miny, l_maxy, l_ni, l_nj, l_nk     ! model (destination) layout
   integer :: pil_w = 0, pil_e = 0, pil_s = 0, pil_n = 0
   integer :: Adz_BC_LAM_flux = 1
   logical :: Adz_BC_LAM_zlf_L = .false.        ! Adz_BC_LAM_flux == 2 (adz_options / adz_set)
   integer :: Schm_advec = 1
   real(kind=REAL64) :: Cstv_dt_8 = 0.d0
   real(kind=REAL64) :: Cstv_bA_m_8 = 0.6d0, Cstv_bA_8 = 0.6d0   ! GEM's default off-centring
   logical :: Adz_BC_min_max_L = .true.
   integer :: Adz_ILMC_sweep_max = 2             ! adz_options.F90:47-53
   logical :: Adz_ILMC_min_max_L = .true.
   logical :: Schm_autobar_L = .false.
   real :: Adz_BC_pexp_list(9) = 1.0
   integer :: Adz_pos_dtype = ADZ_F32       ! storage of positions: ADZ_F32 (the reference's real) / ADZ_F64
   integer :: Adz_field_dtype = ADZ_F32     ! storage of fields (the reference's real)
   integer :: Adz_interp_math = ADZ_MATH_REF
   ! tile parameters of the batched launches (zeros: the library's defaults): adz_main's
   ! groups, and each tracer group with or without post-processing
   type(adz_tune), target :: Adz_tune_main, Adz_tune_tracers
   type(c_ptr) :: Adz_stream = c_null_ptr   ! hipStream_t (null stream by default)
   ! winds read by adz_traject (adz_mem: Adz_uu_ext, Adz_vv_ext, Adz_ww_ext; dynamics: ut0, vt0, zdt0)
   type(c_ptr) :: Adz_uu_ext = c_null_ptr, Adz_vv_ext = c_null_ptr, Adz_ww_ext = c_null_ptr, &
                  ut0 = c_null_ptr, vt0 = c_null_ptr, zdt0 = c_null_ptr
   ! right-hand sides of adz_main (mem_tstp): sources on the extended layout, destinations
   ! on the model layout
   type(c_ptr) :: orhsu_ext = c_null_ptr, orhsv_ext = c_null_ptr, orhsc_ext = c_null_ptr, &
                  orhst_ext = c_null_ptr, orhsf_ext = c_null_ptr, orhsw_ext = c_null_ptr
   type(c_ptr) :: rhsu = c_null_ptr, rhsv = c_null_ptr, rhsc = c_null_ptr, rhst = c_null_ptr, &
                  rhsf = c_null_ptr, rhsw = c_null_ptr
   logical :: Dynamics_hydro_L = .true., Dynamics_hauteur_L = .false.
   ! air mass at P and M, pr_k/pr_s and geomh_mask_8 read by Bermejo_Conde
   type(c_ptr) :: airm1 = c_null_ptr, airm0 = c_null_ptr, pkps = c_null_ptr, geomh_mask_8 = c_null_ptr
   ! with Schm_autobar_L the masses are level-1 sums weighted by geomh_area_mask_8 (geomh)
   type(c_ptr) :: geomh_area_mask_8 = c_null_ptr
   ! inputs of set_post_tr_hlt (gmm_pw, geomh, dyn_fisl_options, tdpack)
   type(c_ptr) :: pw_pm_plus = c_null_ptr, pw_pm_moins = c_null_ptr, pw_pt_moins = c_null_ptr, &
                  pw_p0_moins = c_null_ptr, geomh_area_8 = c_null_ptr
   logical :: Schm_dry_mixing_ratio_L = .false., Schm_wload_L = .false.
   integer :: Tr3d_hu = 0                             ! humidity's index in tracers_P / tracers_M
   integer, allocatable :: Tr3d_wload_idx(:)          ! hydrometeors summed into sumq (tracer order)
   real(kind=REAL64) :: grav_8 = 9.80616d0
   ! tracers (mem_tracers, tr3d): device addresses of element (1,1,1) of each tracer at
   ! time P (source), M (destination) and the ZLF piloting copies B; groups as index ranges
   type(c_ptr), allocatable :: tracers_P(:), tracers_M(:), tracers_B(:)
   integer :: Tr3d_ntrTRICUB_NT = 0, Tr3d_debTRICUB_NT = 1, Tr3d_ntrBICHQV_NT = 0, Tr3d_debBICHQV_NT = 1
   integer :: Tr3d_ntrTRICUB_WP = 0, Tr3d_debTRICUB_WP = 1, Tr3d_ntrBICHQV_WP = 0, Tr3d_debBICHQV_WP = 1
   type(meta_tracers), allocatable, target :: Tr_3CWP(:), Tr_BQWP(:)
   type(Adz_pntr_stack), allocatable :: Adz_stack(:)
   type(post), allocatable, target :: Adz_post_3CWP(:), Adz_post_BQWP(:)
   type(post), pointer :: Adz_post(:) => null()
   type(bc), allocatable :: Adz_bc(:)
   ! masses (mass_fo, mass_fi) of the Aranami flux: one device array for the fixer's tracers
   ! of both groups in registry order; Adz_flux is where adz_tricub_hlt writes the current
   ! group's (adz_tracers points it at Adz_flux_3CWP or Adz_flux_BQWP)
   type(c_ptr) :: Adz_flux = c_null_ptr, Adz_flux_3CWP = c_null_ptr, Adz_flux_BQWP = c_null_ptr
   type(c_ptr), private :: flux_mem = c_null_ptr
   integer, private :: flux_cap = 0
   ! carried deficits of the fixer (device, FP64, one per tracer of the last F_bc)
   type(c_ptr) :: Adz_bc_deficit = c_null_ptr
   integer, private :: deficit_cap = 0
   logical :: Adz_bc_deficit_upload_L = .true.
   integer, private, save :: n_bc = 0, n_bc_TOTAL = 0
   ! set by adz_tracers around its interpolation before the pressure adjustment: the flux
   ! masses are then formed after it, with the air mass of that time (see adz_tracers)
   logical, private, save :: flux_deferred_L = .false.
   logical, private, save :: done_store_BC_L = .false.

   ! ------------------------------------------------------------ plan cache
   integer, parameter, private :: NCACHE = 32
   integer, parameter, private :: K_PLAN = 1, K_BC = 2, K_FLUX = 3, K_ILMC = 4
   type, private :: cached_plan
      integer :: kind = 0
      type(c_ptr) :: h = c_null_ptr
      integer(c_int64_t), allocatable :: key(:)
      integer(c_int64_t) :: used = 0
   end type cached_plan
   type(cached_plan), private, save :: cache(NCACHE)
   integer(c_int64_t), private, save :: cache_clock = 0
   integer(c_int64_t), private, allocatable, save :: kb(:)
   integer, private, save :: kn = 0

contains

   subroutine adz_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc /= ADZ_OK) then
         write (error_unit, '(a,": ",a," (status ",i0,")")') where, adz_error_message(), rc
         error stop 1
      end if
   end subroutine adz_check

   subroutine hip_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc /= 0) then
         write (error_unit, '(a,": HIP error ",i0)') where, rc
         error stop 1
      end if
   end subroutine hip_check

   ! ---------------------------------------------------------------- plan cache helpers
   subroutine key_reset()
      kn = 0
      if (.not. allocated(kb)) allocate (kb(256))
   end subroutine key_reset

   subroutine key_add(v)
      integer(c_int64_t), intent(in) :: v
      integer(c_int64_t), allocatable :: t(:)
      if (kn == size(kb)) then
         allocate (t(2 * size(kb)))
         t(1:kn) = kb(1:kn)
         call move_alloc(t, kb)
      end if
      kn = kn + 1
      kb(kn) = v
   end subroutine key_add

   subroutine key_addi(v)
      integer, intent(in) :: v
      call key_add(int(v, c_int64_t))
   end subroutine key_addi

   subroutine key_addp(p)
      type(c_ptr), intent(in) :: p
      call key_add(int(transfer(p, 0_c_intptr_t), c_int64_t))
   end subroutine key_addp

   ! the cached handle of kind k whose key equals kb(1:kn), or c_null_ptr
   function cache_find(k) result(h)
      integer, intent(in) :: k
      type(c_ptr) :: h
      integer :: s
      h = c_null_ptr
      do s = 1, NCACHE
         if (cache(s)%kind /= k) cycle
         if (size(cache(s)%key) /= kn) cycle
         if (any(cache(s)%key /= kb(1:kn))) cycle
         cache_clock = cache_clock + 1
         cache(s)%used = cache_clock
         h = cache(s)%h
         return
      end do
   end function cache_find

   subroutine cache_release(s)
      integer, intent(in) :: s
      select case (cache(s)%kind)
      case (K_PLAN)
         call adz_check(adz_plan_free(cache(s)%h), 'adz_plan_free')
      case (K_BC)
         call adz_check(adz_bc_free(cache(s)%h), 'adz_bc_free')
      case (K_FLUX)
         call adz_check(adz_flux_free(cache(s)%h), 'adz_flux_free')
      case (K_ILMC)
         call adz_check(adz_ilmc_free(cache(s)%h), 'adz_ilmc_free')
      end select
      cache(s)%kind = 0
      cache(s)%h = c_null_ptr
      if (allocated(cache(s)%key)) deallocate (cache(s)%key)
   end subroutine cache_release

   ! keep handle h of kind k under the key kb(1:kn), releasing the least recently used entry
   subroutine cache_put(k, h)
      integer, intent(in) :: k
      type(c_ptr), intent(in) :: h
      integer :: s, best
      best = 1
      do s = 1, NCACHE
         if (cache(s)%kind == 0) then
            best = s
            exit
         end if
         if (cache(s)%used < cache(best)%used) best = s
      end do
      if (cache(best)%kind /= 0) then
         if (hipDeviceSynchronize() /= 0) call adz_check(ADZ_EHIP, 'plan cache: hipDeviceSynchronize')
         call cache_release(best)
      end if
      cache_clock = cache_clock + 1
      cache(best)%kind = k
      cache(best)%h = h
      cache(best)%key = kb(1:kn)
      cache(best)%used = cache_clock
   end subroutine cache_put

   ! every cached plan and the device arrays of the module (adz_gpu_finalize, or a caller
   ! whose arrays change for good)
   subroutine adz_gpu_release_plans()
      integer :: s
      if (hipDeviceSynchronize() /= 0) call adz_check(ADZ_EHIP, 'adz_gpu_release_plans: hipDeviceSynchronize')
      do s = 1, NCACHE
         if (cache(s)%kind /= 0) call cache_release(s)
      end do
   end subroutine adz_gpu_release_plans

   function c_ptr_offset(p, bytes) result(q)
      type(c_ptr), intent(in) :: p
      integer(c_size_t), intent(in) :: bytes
      type(c_ptr) :: q
      integer(c_intptr_t) :: a
      a = transfer(p, a) + int(bytes, c_intptr_t)
      q = transfer(a, q)
   end function c_ptr_offset

   ! ---------------------------------------------------------------- adz_set
   ! adz_set for one tile: builds the context and fills the module state.
   ! F_model_layout = (l_minx, l_maxx, l_miny, l_maxy) of the model (destination) arrays.
   subroutine adz_gpu_set(F_grid, F_itraj, F_model_layout)
      type(adz_grid), intent(in) :: F_grid
      integer, intent(in) :: F_itraj, F_model_layout(4)
      integer(c_int) :: ints(24), rc
      real(c_double) :: dbls(9)
      integer(c_long) :: ne
      if (c_associated(Adz_ctx)) call adz_gpu_finalize()
      call adz_check(c_adz_set(F_grid, int(Adz_pos_dtype, c_int), int(F_itraj, c_int), Adz_ctx), 'adz_set')
      call adz_check(adz_ctx_info(Adz_ctx, ints, dbls), 'adz_ctx_info')
      Adz_lminx = ints(1); Adz_lmaxx = ints(2); Adz_lminy = ints(3); Adz_lmaxy = ints(4)
      Adz_nit = ints(5); Adz_njt = ints(6); Adz_nij = ints(7)
      Adz_i0 = ints(8); Adz_in = ints(9); Adz_j0 = ints(10); Adz_jn = ints(11)
      Adz_i0u = ints(12); Adz_inu = ints(13); Adz_j0v = ints(14); Adz_jnv = ints(15)
      Adz_i0b = ints(16); Adz_inb = ints(17); Adz_j0b = ints(18); Adz_jnb = ints(19)
      Adz_k0 = ints(20); Adz_k0t = ints(21); Adz_k0m = ints(22)
      l_ni = F_grid%ni; l_nj = F_grid%nj; l_nk = F_grid%nk
      l_minx = F_model_layout(1); l_maxx = F_model_layout(2); l_miny = F_model_layout(3); l_maxy = F_model_layout(4)
      pil_w = merge(F_grid%pil_w, 0, F_grid%west == 1); pil_e = merge(F_grid%pil_e, 0, F_grid%east == 1)
      pil_s = merge(F_grid%pil_s, 0, F_grid%south == 1); pil_n = merge(F_grid%pil_n, 0, F_grid%north == 1)
      Adz_BC_LAM_flux = F_grid%bc_lam_flux
      Adz_BC_LAM_zlf_L = Adz_BC_LAM_flux == 2
      Adz_num_u = (Adz_inu - Adz_i0u + 1) * (Adz_jn - Adz_j0 + 1) * (l_nk - Adz_k0 + 1)
      Adz_num_v = (Adz_in - Adz_i0 + 1) * (Adz_jnv - Adz_j0v + 1) * (l_nk - Adz_k0 + 1)
      Adz_num_q = (Adz_in - Adz_i0 + 1) * (Adz_jn - Adz_j0 + 1) * (l_nk - Adz_k0 + 1)
      Adz_num_t = (Adz_in - Adz_i0 + 1) * (Adz_jn - Adz_j0 + 1) * (l_nk - Adz_k0t + 1)
      Adz_num_b = (Adz_inb - Adz_i0b + 1) * (Adz_jnb - Adz_j0b + 1) * l_nk
      Adz_cpntr_q = adz_ctx_ztab(Adz_ctx, 0_c_int)
      Adz_cpntr_t = adz_ctx_ztab(Adz_ctx, 1_c_int)
      rc = adz_ctx_array(Adz_ctx, ADZ_A_PM, Adz_pm, ne);   call adz_check(rc, 'adz_ctx_array')
      rc = adz_ctx_array(Adz_ctx, ADZ_A_PMU, Adz_pmu, ne); call adz_check(rc, 'adz_ctx_array')
      rc = adz_ctx_array(Adz_ctx, ADZ_A_PMV, Adz_pmv, ne); call adz_check(rc, 'adz_ctx_array')
      rc = adz_ctx_array(Adz_ctx, ADZ_A_PT, Adz_pt, ne);   call adz_check(rc, 'adz_ctx_array')
      rc = adz_ctx_array(Adz_ctx, ADZ_A_PB, Adz_pb, ne);   call adz_check(rc, 'adz_ctx_array')
      rc = adz_ctx_array(Adz_ctx, ADZ_A_PXYZM, Adz_pxyzm, ne); call adz_check(rc, 'adz_ctx_array')
      rc = adz_ctx_array(Adz_ctx, ADZ_A_WPXYZ, Adz_wpxyz, ne); call adz_check(rc, 'adz_ctx_array')
      n_bc = 0; n_bc_TOTAL = 0; done_store_BC_L = .false.
      Adz_bc_deficit_upload_L = .true.
   end subroutine adz_gpu_set

   subroutine adz_gpu_finalize()
      call adz_gpu_release_plans()
      if (c_associated(flux_mem)) call hip_check(hipFree(flux_mem), 'adz_gpu_finalize: hipFree')
      if (c_associated(Adz_bc_deficit)) call hip_check(hipFree(Adz_bc_deficit), 'adz_gpu_finalize: hipFree')
      flux_mem = c_null_ptr; flux_cap = 0
      Adz_bc_deficit = c_null_ptr; deficit_cap = 0
      Adz_flux = c_null_ptr; Adz_flux_3CWP = c_null_ptr; Adz_flux_BQWP = c_null_ptr
      if (c_associated(Adz_ctx)) call adz_check(adz_ctx_free(Adz_ctx), 'adz_ctx_free')
      Adz_ctx = c_null_ptr
   end subroutine adz_gpu_finalize

   subroutine adz_inittraj()
      call adz_check(c_adz_inittraj(Adz_ctx, Adz_stream), 'adz_inittraj')
   end subroutine adz_inittraj

   subroutine adz_traject(F_dt_8)
      real(kind=REAL64), intent(IN) :: F_dt_8
      call adz_check(c_adz_traject(Adz_ctx, F_dt_8, int(Schm_advec, c_int), Cstv_bA_m_8, Cstv_bA_8, Adz_uu_ext, &
                                   Adz_vv_ext, Adz_ww_ext, ut0, vt0, zdt0, Adz_stream), 'adz_traject')
   end subroutine adz_traject

   ! ---------------------------------------------------------------- interpolation
   ! One batched launch over nit items (adz_plan), created on first use and cached under its
   ! field pointers, position sets, boxes, types and tile parameters.
   subroutine run_plan(items, nit, flds, post_flag, tune, where)
      type(adz_item), intent(inout) :: items(:)
      integer, intent(in) :: nit, post_flag
      type(adz_field), intent(in), target :: flds(:)
      type(adz_tune), intent(in), target :: tune
      character(len=*), intent(in) :: where
      type(c_ptr) :: h
      integer :: q, n, f0
      call key_reset()
      call key_addi(nit); call key_addi(post_flag); call key_addi(Adz_field_dtype); call key_addi(Adz_interp_math)
      call key_addi(tune%fields_per_pass); call key_addi(tune%points_outer); call key_addi(tune%blocks_per_cu)
      call key_addi(tune%xcd_chunks); call key_addi(tune%stage); call key_addi(tune%level_inner)
      do q = 1, nit
         call key_addp(items(q)%geom); call key_addp(items(q)%xyz); call key_addi(items(q)%pdtype)
         call key_addi(items(q)%i0); call key_addi(items(q)%in); call key_addi(items(q)%j0); call key_addi(items(q)%jn)
         call key_addi(items(q)%k0); call key_addi(items(q)%nk); call key_addi(items(q)%ext)
         call key_addi(items(q)%dminx); call key_addi(items(q)%dmaxx); call key_addi(items(q)%dminy)
         call key_addi(items(q)%dmaxy); call key_addi(items(q)%nptr)
      end do
      do n = 1, size(flds)
         call key_addp(flds(n)%src); call key_addp(flds(n)%dst); call key_addp(flds(n)%lin)
         call key_addp(flds(n)%mn); call key_addp(flds(n)%mx)
      end do
      h = cache_find(K_PLAN)
      if (.not. c_associated(h)) then
         f0 = 1
         do q = 1, nit  ! each item's field list: its slice of flds (host copies are taken by the plan)
            items(q)%fld = c_loc(flds(f0))
            f0 = f0 + items(q)%nptr
         end do
         call adz_check(adz_plan_create(items, int(nit, c_int), int(Adz_field_dtype, c_int), &
                                        int(Adz_interp_math, c_int), int(post_flag, c_int), c_loc(tune), h), &
                        where//': adz_plan_create')
         call cache_put(K_PLAN, h)
      end if
      call adz_check(adz_plan_execute(h, Adz_stream), where)
   end subroutine run_plan

   subroutine adz_tricub_hlt(F_stk, F_nptr, F_xyz, F_geom, F_num, F_i0, F_in, F_j0, F_jn, F_k0, F_ext_L, F_post, &
                             F_QV_L)
      logical, intent(IN), optional :: F_QV_L, F_ext_L
      integer, intent(in) :: F_nptr, F_num, F_i0, F_in, F_j0, F_jn, F_k0
      type(c_ptr), intent(in) :: F_xyz
      type(meta_tracers), dimension(F_nptr), intent(in), optional :: F_post
      type(Adz_pntr_stack), dimension(F_nptr), intent(inout) :: F_stk
      type(c_ptr), intent(in) :: F_geom
      type(adz_field), allocatable, target :: fl(:)
      type(adz_item) :: it(1)
      integer :: n, pass, cnt, post_flag
      logical :: ext_L, qv_L
      ext_L = .true.
      if (present(F_ext_L)) ext_L = F_ext_L
      qv_L = .false.
      if (present(F_QV_L)) qv_L = F_QV_L
      if (qv_L .and. ext_L) then  ! adz_interp_hlt.F90:63-66
         write (error_unit, *) 'Source must be NOT expended when using quintic - abort'
         error stop 1
      end if
      if (qv_L .and. Adz_field_dtype /= ADZ_F32) then  ! the quintic kernel reads `real` fields
         write (error_unit, *) 'adz_tricub_hlt: F_QV_L needs FP32 fields (Adz_field_dtype = ADZ_F32)'
         error stop 1
      end if
      if ((F_in - F_i0 + 1) <= 0 .or. (F_jn - F_j0 + 1) <= 0 .or. F_nptr <= 0) return
      if (present(F_post) .and. .not. associated(Adz_post)) then
         write (error_unit, *) 'adz_tricub_hlt: F_post needs Adz_post (lin/min/max of the group)'
         error stop 1
      end if
      ! adz_interp_hlt.F90:91-96: the LAM boundary flux of the fixer's tracers
      if (.not. ext_L .and. present(F_post) .and. .not. flux_deferred_L) then
         if (Adz_BC_LAM_flux == 1) call adz_BC_LAM_Aranami_hlt(F_stk, F_post, F_nptr)
      end if
      if (F_num /= (F_in - F_i0 + 1) * (F_jn - F_j0 + 1) * (l_nk - F_k0 + 1)) then
         write (error_unit, *) 'adz_tricub_hlt: F_num does not match the target box'
         error stop 1
      end if
      allocate (fl(F_nptr))
      ! pass 1: fields without post-processing or with mono /= 1; pass 2: mono == 1 (clipped)
      do pass = 1, merge(2, 1, present(F_post))
         cnt = 0
         do n = 1, F_nptr
            if (present(F_post)) then
               if ((F_post(n)%mono == 1) .neqv. (pass == 2)) cycle
            end if
            cnt = cnt + 1
            fl(cnt) = adz_field(F_stk(n)%src, F_stk(n)%dst, c_null_ptr, c_null_ptr, c_null_ptr)
            if (present(F_post)) then
               fl(cnt)%lin = Adz_post(n)%lin
               fl(cnt)%mn = Adz_post(n)%min
               fl(cnt)%mx = Adz_post(n)%max
            end if
         end do
         if (cnt == 0) cycle
         post_flag = 0
         if (present(F_post)) post_flag = ADZ_POST_STORE
         if (pass == 2) post_flag = ior(post_flag, ADZ_POST_CLIP)
         if (qv_L) then  ! bicubHQV_lin_min_max_ntr on the thermodynamic levels (adz_interp_hlt.F90:113-115)
            call adz_check(adz_tricub_hlt_qv(Adz_ctx, fl, int(cnt, c_int), F_xyz, int(F_i0, c_int), &
                                             int(F_in, c_int), int(F_j0, c_int), int(F_jn, c_int), &
                                             int(F_k0, c_int), int(l_minx, c_int), int(l_maxx, c_int), &
                                             int(l_miny, c_int), int(l_maxy, c_int), 0_c_int, &
                                             int(post_flag, c_int), Adz_stream), 'adz_tricub_hlt (F_QV_L)')
            cycle
         end if
         ! all fields of the pass in one persistent launch (the reference's slices of ni*nj
         ! points on OpenMP threads, adz_interp_hlt.F90:102-107)
         it(1) = adz_item(F_geom, F_xyz, int(Adz_pos_dtype, c_int), F_i0, F_in, F_j0, F_jn, F_k0, l_nk, &
                          l_minx, l_maxx, l_miny, l_maxy, merge(1, 0, ext_L), c_null_ptr, cnt)
         call run_plan(it, 1, fl(1:cnt), post_flag, Adz_tune_tracers, 'adz_tricub_hlt')
      end do
      deallocate (fl)
   end subroutine adz_tricub_hlt

   ! adz_main.F90:16-71: the trajectory, then u on Adz_pmu, v on Adz_pmv, rhsc on Adz_pm
   ! (momentum levels) and rhst, rhsf (and rhsw when non-hydrostatic or in height
   ! coordinates) on Adz_pt (thermodynamic levels), sources on the extended layout.  The four
   ! groups run as one persistent launch.  The halo exchange of rhsu/rhsv that ends the
   ! reference routine moves nothing on one tile.
   subroutine adz_main()
      type(adz_item) :: it(4)
      type(adz_field), target :: fl(6)
      integer :: n
      call adz_traject(Cstv_dt_8)
      fl(1) = adz_field(orhsu_ext, rhsu, c_null_ptr, c_null_ptr, c_null_ptr)
      fl(2) = adz_field(orhsv_ext, rhsv, c_null_ptr, c_null_ptr, c_null_ptr)
      fl(3) = adz_field(orhsc_ext, rhsc, c_null_ptr, c_null_ptr, c_null_ptr)
      fl(4) = adz_field(orhst_ext, rhst, c_null_ptr, c_null_ptr, c_null_ptr)
      fl(5) = adz_field(orhsf_ext, rhsf, c_null_ptr, c_null_ptr, c_null_ptr)
      n = 2
      if ((.not. Dynamics_hydro_L) .or. Dynamics_hauteur_L) then
         n = 3
         fl(6) = adz_field(orhsw_ext, rhsw, c_null_ptr, c_null_ptr, c_null_ptr)
      end if
      it(1) = adz_item(Adz_cpntr_q, Adz_pmu, int(Adz_pos_dtype, c_int), Adz_i0u, Adz_inu, Adz_j0, Adz_jn, Adz_k0, &
                       l_nk, l_minx, l_maxx, l_miny, l_maxy, 1, c_null_ptr, 1)
      it(2) = adz_item(Adz_cpntr_q, Adz_pmv, int(Adz_pos_dtype, c_int), Adz_i0, Adz_in, Adz_j0v, Adz_jnv, Adz_k0, &
                       l_nk, l_minx, l_maxx, l_miny, l_maxy, 1, c_null_ptr, 1)
      it(3) = adz_item(Adz_cpntr_q, Adz_pm, int(Adz_pos_dtype, c_int), Adz_i0, Adz_in, Adz_j0, Adz_jn, Adz_k0, &
                       l_nk, l_minx, l_maxx, l_miny, l_maxy, 1, c_null_ptr, 1)
      it(4) = adz_item(Adz_cpntr_t, Adz_pt, int(Adz_pos_dtype, c_int), Adz_i0, Adz_in, Adz_j0, Adz_jn, Adz_k0t, &
                       l_nk, l_minx, l_maxx, l_miny, l_maxy, 1, c_null_ptr, n)
      call run_plan(it, 4, fl(1:3 + n), 0, Adz_tune_main, 'adz_main')
   end subroutine adz_main

   ! set_post_tr_hlt.F90:18-70: the air masses at P (airm1) and M (airm0) and pr_t/pr_s at M
   ! (pkps), from the pressures of gmm_pw, the humidity (with Schm_dry_mixing_ratio_L) and
   ! the hydrometeors (with Schm_wload_L as well), on the module's arrays.
   subroutine set_post_tr_hlt()
      type(adz_airmass_fields) :: f
      type(c_ptr), allocatable, target :: hp(:), hm(:)
      integer :: n, nh
      f%minx = l_minx; f%maxx = l_maxx; f%miny = l_miny; f%maxy = l_maxy
      f%pm_p = pw_pm_plus; f%pm_m = pw_pm_moins; f%pt_m = pw_pt_moins; f%p0_m = pw_p0_moins
      f%area = geomh_area_8
      f%dry_mixing_ratio = merge(1, 0, Schm_dry_mixing_ratio_L)
      f%wload = merge(1, 0, Schm_wload_L)
      if (Schm_dry_mixing_ratio_L) then
         if (Tr3d_hu < 1) then
            write (error_unit, *) 'set_post_tr_hlt: Schm_dry_mixing_ratio_L needs Tr3d_hu'
            error stop 1
         end if
         f%hu_p = tracers_P(Tr3d_hu); f%hu_m = tracers_M(Tr3d_hu)
      end if
      nh = 0
      if (allocated(Tr3d_wload_idx)) nh = size(Tr3d_wload_idx)
      allocate (hp(max(nh, 1)), hm(max(nh, 1)))
      do n = 1, nh
         hp(n) = tracers_P(Tr3d_wload_idx(n)); hm(n) = tracers_M(Tr3d_wload_idx(n))
      end do
      f%nhyd = nh
      if (nh > 0) then
         f%hyd_p = c_loc(hp); f%hyd_m = c_loc(hm)
      end if
      call adz_check(adz_set_post_tr(Adz_ctx, f, grav_8, airm1, airm0, pkps, int(Adz_field_dtype, c_int), &
                                     Adz_stream), 'set_post_tr_hlt')
      deallocate (hp, hm)
   end subroutine set_post_tr_hlt

   ! point Adz_stack at tracers deb .. deb+n-1 of tracers_P (src) / tracers_M (dst)
   subroutine stack_group(deb, n, with_pil)
      integer, intent(in) :: deb, n
      logical, intent(in) :: with_pil
      type(Adz_pntr_stack), allocatable :: t(:)
      integer :: q
      if (.not. allocated(Adz_stack)) allocate (Adz_stack(max(n, 1)))
      if (size(Adz_stack) < n) then
         allocate (t(n))
         call move_alloc(t, Adz_stack)
      end if
      do q = 1, n
         Adz_stack(q)%src = tracers_P(deb + q - 1)
         Adz_stack(q)%dst = tracers_M(deb + q - 1)
         if (with_pil) Adz_stack(q)%pil = tracers_B(deb + q - 1)
      end do
   end subroutine stack_group

   ! adz_tracers.F90:16-184.  Before the pressure adjustment (F_before_psadj_L): the
   ! interpolation of every group (tricubic / quintic, without and with post-processing);
   ! after it: the air masses, ILMC, and the fixer on the tracers registered for it.
   ! The Aranami flux of a group is formed after the adjustment, from the group's sources
   ! (tracers_P) at Adz_pb, neither of which the adjustment changes: the reference
   ! interpolates FLUX_out / FLUX_in in adz_tricub_hlt and sums their masses in
   ! Bermejo_Conde with the air mass of time M that set_post_tr_hlt forms after the
   ! adjustment (Bermejo_Conde.F90:77-82); here both happen at that point.
   subroutine adz_tracers(F_before_psadj_L)
      logical, intent(in) :: F_before_psadj_L
      logical :: after_psadj_L
      after_psadj_L = .not. F_before_psadj_L
      if (Tr3d_ntrTRICUB_NT > 0 .and. F_before_psadj_L) then
         call stack_group(Tr3d_debTRICUB_NT, Tr3d_ntrTRICUB_NT, .false.)
         call adz_tricub_hlt(Adz_stack, Tr3d_ntrTRICUB_NT, Adz_pt, Adz_cpntr_t, Adz_num_t, Adz_i0, Adz_in, &
                             Adz_j0, Adz_jn, Adz_k0t, F_ext_L=.false.)
      end if
      if (Tr3d_ntrBICHQV_NT > 0 .and. F_before_psadj_L) then
         call stack_group(Tr3d_debBICHQV_NT, Tr3d_ntrBICHQV_NT, .false.)
         call adz_tricub_hlt(Adz_stack, Tr3d_ntrBICHQV_NT, Adz_pt, Adz_cpntr_t, Adz_num_t, Adz_i0, Adz_in, &
                             Adz_j0, Adz_jn, Adz_k0t, F_ext_L=.false., F_QV_L=.true.)
      end if
      if (max(Tr3d_ntrTRICUB_WP, Tr3d_ntrBICHQV_WP) > 0 .and. after_psadj_L) call set_post_tr_hlt()
      if (Tr3d_ntrTRICUB_WP > 0 .or. Tr3d_ntrBICHQV_WP > 0) call flux_masses_layout()
      if (Tr3d_ntrTRICUB_WP > 0) then
         Adz_post => Adz_post_3CWP
         Adz_flux = Adz_flux_3CWP
         call stack_group(Tr3d_debTRICUB_WP, Tr3d_ntrTRICUB_WP, .false.)
         if (.not. Adz_BC_LAM_zlf_L .and. F_before_psadj_L) then
            flux_deferred_L = .true.
            call adz_tricub_hlt(Adz_stack, Tr3d_ntrTRICUB_WP, Adz_pt, Adz_cpntr_t, Adz_num_t, Adz_i0, Adz_in, &
                                Adz_j0, Adz_jn, Adz_k0t, F_ext_L=.false., F_post=Tr_3CWP)
            flux_deferred_L = .false.
         else if (F_before_psadj_L) then
            call stack_group(Tr3d_debTRICUB_WP, Tr3d_ntrTRICUB_WP, .true.)
            call adz_BC_LAM_zlf_0_hlt(Tr3d_ntrTRICUB_WP, 0)
            call adz_tricub_hlt(Adz_stack, Tr3d_ntrTRICUB_WP, Adz_pb, Adz_cpntr_t, Adz_num_b, Adz_i0b, Adz_inb, &
                                Adz_j0b, Adz_jnb, 1, F_ext_L=.false., F_post=Tr_3CWP)
            call adz_BC_LAM_zlf_0_hlt(Tr3d_ntrTRICUB_WP, 1)
         end if
         if (after_psadj_L) then
            if (Adz_BC_LAM_flux == 1) call adz_BC_LAM_Aranami_hlt(Adz_stack, Tr_3CWP, Tr3d_ntrTRICUB_WP)
            call adz_post_tr(1)
         end if
      end if
      if (Tr3d_ntrBICHQV_WP > 0) then
         Adz_post => Adz_post_BQWP
         Adz_flux = Adz_flux_BQWP
         call stack_group(Tr3d_debBICHQV_WP, Tr3d_ntrBICHQV_WP, .false.)
         if (.not. Adz_BC_LAM_zlf_L .and. F_before_psadj_L) then
            flux_deferred_L = .true.
            call adz_tricub_hlt(Adz_stack, Tr3d_ntrBICHQV_WP, Adz_pt, Adz_cpntr_t, Adz_num_t, Adz_i0, Adz_in, &
                                Adz_j0, Adz_jn, Adz_k0t, F_ext_L=.false., F_QV_L=.true., F_post=Tr_BQWP)
            flux_deferred_L = .false.
         else if (F_before_psadj_L) then
            call stack_group(Tr3d_debBICHQV_WP, Tr3d_ntrBICHQV_WP, .true.)
            call adz_BC_LAM_zlf_0_hlt(Tr3d_ntrBICHQV_WP, 0)
            call adz_tricub_hlt(Adz_stack, Tr3d_ntrBICHQV_WP, Adz_pb, Adz_cpntr_t, Adz_num_b, Adz_i0b, Adz_inb, &
                                Adz_j0b, Adz_jnb, 1, F_ext_L=.false., F_QV_L=.true., F_post=Tr_BQWP)
            call adz_BC_LAM_zlf_0_hlt(Tr3d_ntrBICHQV_WP, 1)
         end if
         if (after_psadj_L) then
            if (Adz_BC_LAM_flux == 1) call adz_BC_LAM_Aranami_hlt(Adz_stack, Tr_BQWP, Tr3d_ntrBICHQV_WP)
            call adz_post_tr(2)
         end if
      end if
      if (max(Tr3d_ntrTRICUB_WP, Tr3d_ntrBICHQV_WP) > 0 .and. after_psadj_L) call adz_post_tr(0)
      if (Tr3d_ntrTRICUB_WP > 0 .and. Adz_BC_LAM_zlf_L .and. after_psadj_L) then
         call stack_group(Tr3d_debTRICUB_WP, Tr3d_ntrTRICUB_WP, .true.)
         call adz_BC_LAM_zlf_0_hlt(Tr3d_ntrTRICUB_WP, 2)
      end if
      if (Tr3d_ntrBICHQV_WP > 0 .and. Adz_BC_LAM_zlf_L .and. after_psadj_L) then
         call stack_group(Tr3d_debBICHQV_WP, Tr3d_ntrBICHQV_WP, .true.)
         call adz_BC_LAM_zlf_0_hlt(Tr3d_ntrBICHQV_WP, 2)
      end if
   end subroutine adz_tracers

   ! ---------------------------------------------------------------- LAM boundary flux
   ! the weight of the mass sums: geomh_mask_8, or geomh_area_mask_8 under Schm_autobar_L
   ! (mass_tr_PM.F90:79-85, mass_tr_X.F90:73-79)
   function mass_mask() result(p)
      type(c_ptr) :: p
      p = geomh_mask_8
      if (Schm_autobar_L) then
         if (.not. c_associated(geomh_area_mask_8)) then
            write (error_unit, *) 'Schm_autobar_L: geomh_area_mask_8 must be set before the tracers'
            error stop 1
         end if
         p = geomh_area_mask_8
      end if
   end function mass_mask

   logical function uses_bc(m)
      type(meta_tracers), intent(in) :: m
      uses_bc = m%mass == 1 .or. (m%mass >= 111 .and. m%mass <= 139)
   end function uses_bc

   ! The device array of flux masses: the fixer's tracers of Tr_3CWP(1:Tr3d_ntrTRICUB_WP)
   ! then of Tr_BQWP(1:Tr3d_ntrBICHQV_WP), (mass_fo, mass_fi) each: the order in which
   ! adz_post_tr registers them (Adz_bc).
   subroutine flux_masses_layout()
      integer :: n1, n2, n
      n1 = 0; n2 = 0
      do n = 1, Tr3d_ntrTRICUB_WP
         if (uses_bc(Tr_3CWP(n))) n1 = n1 + 1
      end do
      do n = 1, Tr3d_ntrBICHQV_WP
         if (uses_bc(Tr_BQWP(n))) n2 = n2 + 1
      end do
      call ensure_flux_mem(n1 + n2)
      Adz_flux_3CWP = flux_mem
      Adz_flux_BQWP = c_ptr_offset(flux_mem, 16_c_size_t * int(n1, c_size_t))
   end subroutine flux_masses_layout

   subroutine ensure_flux_mem(n)
      integer, intent(in) :: n
      if (n <= flux_cap .and. c_associated(flux_mem)) return
      if (c_associated(flux_mem)) then
         if (hipDeviceSynchronize() /= 0) call adz_check(ADZ_EHIP, 'flux masses: hipDeviceSynchronize')
         call hip_check(hipFree(flux_mem), 'flux masses: hipFree')
      end if
      flux_cap = max(n, 1)
      call hip_check(hipMalloc(flux_mem, 16_c_size_t * int(flux_cap, c_size_t)), 'flux masses: hipMalloc')
   end subroutine ensure_flux_mem

   ! adz_BC_LAM_Aranami_hlt.F90:18-236 for the tracers of F_post that use the fixer (mass 1
   ! or 111..139): the masses of FLUX_out / FLUX_in at Adz_pb, written at Adz_flux in the
   ! order of those tracers (no host synchronisation).
   subroutine adz_BC_LAM_Aranami_hlt(F_stk, F_post, F_nptr)
      integer, intent(in) :: F_nptr
      type(Adz_pntr_stack), dimension(F_nptr), intent(in) :: F_stk
      type(meta_tracers), dimension(F_nptr), intent(in) :: F_post
      type(adz_flux_tracer), allocatable :: tr(:)
      type(adz_flux_domain) :: dom
      type(c_ptr) :: h
      integer :: n, nf
      allocate (tr(F_nptr))
      nf = 0
      do n = 1, F_nptr
         if (.not. uses_bc(F_post(n))) cycle
         nf = nf + 1
         tr(nf)%src = F_stk(n)%src
      end do
      if (nf == 0) return
      if (.not. (c_associated(airm0) .and. c_associated(geomh_mask_8))) then
         write (error_unit, *) 'adz_BC_LAM_Aranami_hlt: airm0 and geomh_mask_8 must be set before the tracers'
         error stop 1
      end if
      if (.not. c_associated(Adz_flux)) then  ! a caller outside adz_tracers: one group, at the start
         call ensure_flux_mem(nf)
         Adz_flux = flux_mem
         Adz_flux_3CWP = flux_mem
      end if
      dom = adz_flux_domain(l_minx, l_maxx, l_miny, l_maxy, l_nk, l_ni, l_nj, Adz_i0b, Adz_inb, Adz_j0b, Adz_jnb, &
                            1 + pil_w, l_ni - pil_e, 1 + pil_s, l_nj - pil_n, Adz_k0t, merge(1, 0, Schm_autobar_L))
      call key_reset()
      call key_addi(nf); call key_addi(Adz_field_dtype); call key_addi(Adz_pos_dtype); call key_addp(Adz_cpntr_t)
      call key_addi(l_minx); call key_addi(l_maxx); call key_addi(l_miny); call key_addi(l_maxy); call key_addi(l_nk)
      call key_addi(Adz_i0b); call key_addi(Adz_inb); call key_addi(Adz_j0b); call key_addi(Adz_jnb)
      call key_addi(pil_w); call key_addi(pil_e); call key_addi(pil_s); call key_addi(pil_n); call key_addi(Adz_k0t)
      call key_addi(merge(1, 0, Schm_autobar_L))
      do n = 1, nf
         call key_addp(tr(n)%src)
      end do
      h = cache_find(K_FLUX)
      if (.not. c_associated(h)) then
         call adz_check(adz_flux_create(tr, int(nf, c_int), dom, Adz_cpntr_t, int(Adz_field_dtype, c_int), &
                                        int(Adz_pos_dtype, c_int), h), 'adz_flux_create')
         call cache_put(K_FLUX, h)
      end if
      call adz_check(c_adz_bc_lam_flux(h, Adz_pb, airm0, mass_mask(), Adz_flux, Adz_stream), 'adz_bc_lam_flux')
      deallocate (tr)
   end subroutine adz_BC_LAM_Aranami_hlt

   ! adz_BC_LAM_zlf_hlt (adz_BC_LAM_zlf_hlt.F90:18-319) on n device arrays (element (1,1,1)
   ! of the model layout); F2 is read or written by setups 2 and 4 only.
   subroutine adz_BC_LAM_zlf_hlt(F1, F2, F_n, F_setup)
      integer, intent(in) :: F_n, F_setup
      type(c_ptr), intent(in) :: F1(F_n), F2(F_n)
      type(adz_zlf_domain) :: dom
      if (F_n <= 0) return
      dom = adz_zlf_domain(l_minx, l_maxx, l_miny, l_maxy, l_nk, l_ni, l_nj, pil_w, pil_e, pil_s, pil_n, &
                           1 - Adz_lminx, Adz_k0t)
      call adz_check(adz_bc_lam_zlf(int(F_setup, c_int), F1, F2, int(F_n, c_int), dom, int(Adz_field_dtype, c_int), &
                                    Adz_stream), 'adz_BC_LAM_zlf_hlt')
   end subroutine adz_BC_LAM_zlf_hlt

   ! adz_BC_LAM_zlf_0_hlt (adz_BC_LAM_zlf_0_hlt.F90:18-102) on Adz_stack(1:F_nptr):
   ! F_number 0 before the interpolation (setup 1 on src, setup 2 from dst into pil),
   ! 1 before adz_post_tr (setup 3 on dst), 2 after it (setup 4 from pil into dst).
   subroutine adz_BC_LAM_zlf_0_hlt(F_nptr, F_number)
      integer, intent(in) :: F_nptr, F_number
      type(c_ptr) :: s(F_nptr), d(F_nptr), p(F_nptr)
      integer :: n
      do n = 1, F_nptr
         s(n) = Adz_stack(n)%src; d(n) = Adz_stack(n)%dst; p(n) = Adz_stack(n)%pil
      end do
      if (F_number == 0) then
         call adz_BC_LAM_zlf_hlt(s, s, F_nptr, 1)
         call adz_BC_LAM_zlf_hlt(d, p, F_nptr, 2)
      else if (F_number == 1) then
         call adz_BC_LAM_zlf_hlt(d, d, F_nptr, 3)
      else if (F_number == 2) then
         call adz_BC_LAM_zlf_hlt(d, p, F_nptr, 4)
      end if
   end subroutine adz_BC_LAM_zlf_0_hlt

   ! ---------------------------------------------------------------- post-processing
   ! adz_post_tr.F90:18-276.  F_wp = 1 (Tr_3CWP) or 2 (Tr_BQWP): ILMC on the group's tracers
   ! whose mono code is 2 (:104; the clipping of mono = 1 happened in adz_tricub_hlt), and,
   ! on the first step, the registry of the tracers that use the fixer (:108-117);
   ! F_wp = 0: core limits, settings from each tracer's mass code, Bermejo_Conde on the
   ! registry (:139-228).
   subroutine adz_post_tr(F_wp)
      integer, intent(in) :: F_wp
      type(meta_tracers), dimension(:), pointer :: F_post
      type(bc), allocatable :: t(:)
      integer, allocatable :: il(:)
      integer :: n, num, i0, in, j0, jn, k0, F_nptr, wp, deb, nil
      if (F_wp /= 0) then
         if (F_wp == 1) then
            if (.not. allocated(Tr_3CWP)) return
            F_post => Tr_3CWP; F_nptr = Tr3d_ntrTRICUB_WP
         else
            if (.not. allocated(Tr_BQWP)) return
            F_post => Tr_BQWP; F_nptr = Tr3d_ntrBICHQV_WP
         end if
         if (F_nptr == 0) F_nptr = size(F_post)
         allocate (il(F_nptr))
         nil = 0
         do n = 1, F_nptr
            if (F_post(n)%mono == 2) then
               nil = nil + 1
               il(nil) = n
            end if
         end do
         if (nil > 0) call ilmc_tracers(il(1:nil), Adz_i0, Adz_in, Adz_j0, Adz_jn, 'TR/ group')
         deallocate (il)
         if (done_store_BC_L) return
         if (.not. allocated(Adz_bc)) allocate (Adz_bc(0))
         do n = 1, F_nptr
            if (.not. uses_bc(F_post(n))) cycle
            n_bc = n_bc + 1
            if (size(Adz_bc) < n_bc) then
               allocate (t(max(2 * size(Adz_bc), n_bc)))
               t(1:size(Adz_bc)) = Adz_bc
               call move_alloc(t, Adz_bc)
            end if
            Adz_bc(n_bc)%n = n
            Adz_bc(n_bc)%wp = F_wp
         end do
         n_bc_TOTAL = n_bc
         return
      end if
      done_store_BC_L = .true.
      if (n_bc_TOTAL == 0) return
      i0 = 1 + pil_w; in = l_ni - pil_e; j0 = 1 + pil_s; jn = l_nj - pil_n; k0 = Adz_k0t
      if (Adz_BC_LAM_flux == 2) then  ! ZLF: the fixer's scope is the Adz_pb box (adz_post_tr.F90:147-151)
         i0 = Adz_i0b; in = Adz_inb; j0 = Adz_j0b; jn = Adz_jnb; k0 = 1
      end if
      do num = 1, n_bc_TOTAL
         n = Adz_bc(num)%n
         wp = Adz_bc(num)%wp
         if (wp == 1) then
            F_post => Tr_3CWP; Adz_post => Adz_post_3CWP; deb = Tr3d_debTRICUB_WP
         else
            F_post => Tr_BQWP; Adz_post => Adz_post_BQWP; deb = Tr3d_debBICHQV_WP
         end if
         if (allocated(tracers_P)) then
            Adz_bc(num)%p = tracers_P(deb + n - 1)
            Adz_bc(num)%m = tracers_M(deb + n - 1)
         else  ! a caller that fills Adz_stack itself (one group)
            Adz_bc(num)%p = Adz_stack(n)%src
            Adz_bc(num)%m = Adz_stack(n)%dst
         end if
         Adz_bc(num)%lin = Adz_post(n)%lin
         Adz_bc(num)%min = Adz_post(n)%min
         Adz_bc(num)%max = Adz_post(n)%max
         Adz_bc(num)%name_S = 'TR/'//trim(F_post(n)%name)//':M'
         Adz_bc(num)%mass_deficit = F_post(n)%BC_mass_deficit
         Adz_bc(num)%LEGACY_L = F_post(n)%mass == 1
         Adz_bc(num)%weight = 1
         Adz_bc(num)%pexp_n = 0
         if (.not. Adz_bc(num)%LEGACY_L) then
            Adz_bc(num)%weight = (F_post(n)%mass - 100) / 10
            Adz_bc(num)%pexp_n = F_post(n)%mass - (100 + Adz_bc(num)%weight * 10)
         end if
      end do
      call Bermejo_Conde(Adz_bc, n_bc_TOTAL, i0, in, j0, jn, k0)
   end subroutine adz_post_tr

   ! The carried deficits (device) back into Adz_bc(:)%mass_deficit and the tracers' metas
   ! (adz_post_tr.F90:218-228): for output or a restart file.  Synchronises Adz_stream.
   subroutine adz_bc_sync_deficits()
      real(kind=REAL64), allocatable, target :: h(:)
      integer :: num
      if (n_bc_TOTAL == 0 .or. .not. c_associated(Adz_bc_deficit)) return
      allocate (h(n_bc_TOTAL))
      call hip_check(hipMemcpyAsync(c_loc(h), Adz_bc_deficit, 8_c_size_t * int(n_bc_TOTAL, c_size_t), &
                                    hipMemcpyDeviceToHost, Adz_stream), 'adz_bc_sync_deficits: hipMemcpyAsync')
      call hip_check(hipStreamSynchronize(Adz_stream), 'adz_bc_sync_deficits: hipStreamSynchronize')
      do num = 1, n_bc_TOTAL
         Adz_bc(num)%mass_deficit = h(num)
         if (Adz_bc(num)%wp == 2) then
            Tr_BQWP(Adz_bc(num)%n)%BC_mass_deficit = h(num)
         else if (allocated(Tr_3CWP)) then
            Tr_3CWP(Adz_bc(num)%n)%BC_mass_deficit = h(num)
         end if
      end do
      deallocate (h)
   end subroutine adz_bc_sync_deficits

   ! The metas' BC_mass_deficit into the device deficits of the registry (after a restart
   ! file is read, or to start again): synchronous, outside any capture.
   subroutine adz_bc_upload_deficits()
      real(kind=REAL64), allocatable, target :: h(:)
      integer :: num
      if (n_bc_TOTAL == 0 .or. .not. c_associated(Adz_bc_deficit)) then
         Adz_bc_deficit_upload_L = .true.  ! the first Bermejo_Conde call reads the metas
         return
      end if
      allocate (h(n_bc_TOTAL))
      do num = 1, n_bc_TOTAL
         if (Adz_bc(num)%wp == 2) then
            h(num) = Tr_BQWP(Adz_bc(num)%n)%BC_mass_deficit
         else
            h(num) = Tr_3CWP(Adz_bc(num)%n)%BC_mass_deficit
         end if
         Adz_bc(num)%mass_deficit = h(num)
      end do
      call hip_check(hipMemcpyAsync(Adz_bc_deficit, c_loc(h), 8_c_size_t * int(n_bc_TOTAL, c_size_t), &
                                    hipMemcpyHostToDevice, Adz_stream), 'adz_bc_upload_deficits: hipMemcpyAsync')
      call hip_check(hipStreamSynchronize(Adz_stream), 'adz_bc_upload_deficits: hipStreamSynchronize')
      Adz_bc_deficit_upload_L = .false.
      deallocate (h)
   end subroutine adz_bc_upload_deficits

   ! ILMC_LAM.F90:18-273 on Adz_stack(F_ns)%dst with Adz_post(F_ns)%min / %max and the air
   ! mass at time M (airm0), over F_i0:F_in x F_j0:F_jn x Adz_k0t:l_nk of a tile that touches
   ! the four LAM edges.  F_name_S names the tracer in messages.
   subroutine ILMC_LAM(F_ns, F_name_S, F_i0, F_in, F_j0, F_jn)
      character(len=*), intent(in) :: F_name_S
      integer, intent(in) :: F_ns, F_i0, F_in, F_j0, F_jn
      call ilmc_tracers([F_ns], F_i0, F_in, F_j0, F_jn, F_name_S)
   end subroutine ILMC_LAM

   ! ILMC on the tracers ns(:) of the current group at once (they do not interact): one
   ! cached plan, one launch per hyperplane for all of them
   subroutine ilmc_tracers(ns, F_i0, F_in, F_j0, F_jn, F_name_S)
      integer, intent(in) :: ns(:), F_i0, F_in, F_j0, F_jn
      character(len=*), intent(in) :: F_name_S
      type(adz_ilmc_tracer), allocatable :: tr(:)
      type(adz_ilmc_domain) :: dom
      type(c_ptr) :: h
      integer :: q
      if (Adz_field_dtype /= ADZ_F32) then  ! the ILMC kernels read and write `real` arrays
         write (error_unit, *) 'ILMC_LAM: needs FP32 fields (Adz_field_dtype = ADZ_F32) for ', trim(F_name_S)
         error stop 1
      end if
      if (.not. c_associated(airm0)) then
         write (error_unit, *) 'ILMC_LAM: airm0 must be set before ', trim(F_name_S)
         error stop 1
      end if
      if (.not. associated(Adz_post)) then
         write (error_unit, *) 'ILMC_LAM: Adz_post (min/max of the group) must be set before ', trim(F_name_S)
         error stop 1
      end if
      allocate (tr(size(ns)))
      do q = 1, size(ns)
         tr(q) = adz_ilmc_tracer(Adz_stack(ns(q))%dst, Adz_post(ns(q))%min, Adz_post(ns(q))%max)
      end do
      dom = adz_ilmc_domain(l_minx, l_maxx, l_miny, l_maxy, l_nk, F_i0, F_in, F_j0, F_jn, Adz_k0t, &
                            Adz_ILMC_sweep_max, merge(1, 0, Adz_ILMC_min_max_L), merge(1, 0, Schm_autobar_L))
      call key_reset()
      call key_addi(size(ns)); call key_addi(l_minx); call key_addi(l_maxx); call key_addi(l_miny)
      call key_addi(l_maxy); call key_addi(l_nk); call key_addi(F_i0); call key_addi(F_in); call key_addi(F_j0)
      call key_addi(F_jn); call key_addi(Adz_k0t); call key_addi(Adz_ILMC_sweep_max)
      call key_addi(merge(1, 0, Adz_ILMC_min_max_L)); call key_addi(merge(1, 0, Schm_autobar_L))
      do q = 1, size(ns)
         call key_addp(tr(q)%f); call key_addp(tr(q)%mn); call key_addp(tr(q)%mx)
      end do
      h = cache_find(K_ILMC)
      if (.not. c_associated(h)) then
         call adz_check(adz_ilmc_create(tr, int(size(ns), c_int), dom, h), &
                        'ILMC_LAM ('//trim(F_name_S)//'): adz_ilmc_create')
         call cache_put(K_ILMC, h)
      end if
      call adz_check(adz_ilmc_lam(h, airm0, c_null_ptr, Adz_stream), 'ILMC_LAM ('//trim(F_name_S)//')')
      deallocate (tr)
   end subroutine ilmc_tracers

   ! Bermejo_Conde.F90:18-320 on the F_ntr_bc tracers of F_bc over F_i0:F_in x F_j0:F_jn x
   ! F_k0:l_nk, with the module's air masses, pr_k/pr_s and mask, and, with
   ! Adz_BC_LAM_flux = 1, the flux masses adz_tricub_hlt wrote for the registry's tracers.
   ! The carried deficits stay in Adz_bc_deficit (device): F_bc(:)%mass_deficit is read when
   ! Adz_bc_deficit_upload_L is set (the first call, a restart) and written back by
   ! adz_bc_sync_deficits.
   subroutine Bermejo_Conde(F_bc, F_ntr_bc, F_i0, F_in, F_j0, F_jn, F_k0)
      integer, intent(in) :: F_i0, F_in, F_j0, F_jn, F_k0, F_ntr_bc
      type(bc), intent(inout), dimension(F_ntr_bc) :: F_bc
      type(adz_bc_tracer), allocatable :: tr(:)
      type(adz_bc_domain) :: dom
      type(c_ptr) :: h
      real(kind=REAL64), allocatable, target :: def(:)
      integer :: n, n1, q
      logical :: lam_flux_L
      if (F_ntr_bc <= 0) return
      lam_flux_L = Adz_BC_LAM_flux == 1 .and. c_associated(flux_mem)
      if (lam_flux_L) then
         ! the flux masses follow the registry: the fixer's tracers of Tr_3CWP, then of Tr_BQWP
         n1 = 0
         if (allocated(Tr_3CWP) .and. Tr3d_ntrTRICUB_WP > 0) then
            do n = 1, Tr3d_ntrTRICUB_WP
               if (uses_bc(Tr_3CWP(n))) n1 = n1 + 1
            end do
         end if
         q = count(F_bc(:)%wp == 2)
         if (q > 0 .and. (F_ntr_bc - q /= n1 .or. any(F_bc(n1 + 1:F_ntr_bc)%wp /= 2))) then
            write (error_unit, *) 'Bermejo_Conde: F_bc must list the fixer''s tracers in registry order'
            error stop 1
         end if
      end if
      allocate (tr(F_ntr_bc))
      do n = 1, F_ntr_bc
         tr(n) = adz_bc_tracer(F_bc(n)%p, F_bc(n)%m, F_bc(n)%lin, F_bc(n)%min, F_bc(n)%max, &
                               merge(1, 0, F_bc(n)%LEGACY_L), F_bc(n)%weight, F_bc(n)%pexp_n)
      end do
      dom = adz_bc_domain(l_minx, l_maxx, l_miny, l_maxy, l_nk, F_i0, F_in, F_j0, F_jn, F_k0, &
                          merge(1, 0, Adz_BC_min_max_L), Adz_BC_pexp_list, merge(1, 0, Schm_autobar_L))
      ! the device deficits: one per tracer of F_bc, carried from call to call
      if (F_ntr_bc > deficit_cap .or. .not. c_associated(Adz_bc_deficit)) then
         if (c_associated(Adz_bc_deficit)) then
            if (hipDeviceSynchronize() /= 0) call adz_check(ADZ_EHIP, 'Bermejo_Conde: hipDeviceSynchronize')
            call hip_check(hipFree(Adz_bc_deficit), 'Bermejo_Conde: hipFree')
         end if
         deficit_cap = F_ntr_bc
         call hip_check(hipMalloc(Adz_bc_deficit, 8_c_size_t * int(deficit_cap, c_size_t)), &
                        'Bermejo_Conde: hipMalloc')
         Adz_bc_deficit_upload_L = .true.
      end if
      if (Adz_bc_deficit_upload_L) then
         allocate (def(F_ntr_bc))
         do n = 1, F_ntr_bc
            def(n) = F_bc(n)%mass_deficit
         end do
         call hip_check(hipMemcpyAsync(Adz_bc_deficit, c_loc(def), 8_c_size_t * int(F_ntr_bc, c_size_t), &
                                       hipMemcpyHostToDevice, Adz_stream), 'Bermejo_Conde: hipMemcpyAsync')
         call hip_check(hipStreamSynchronize(Adz_stream), 'Bermejo_Conde: hipStreamSynchronize')
         deallocate (def)
         Adz_bc_deficit_upload_L = .false.
      end if
      call key_reset()
      call key_addi(F_ntr_bc); call key_addi(Adz_field_dtype)
      call key_addi(l_minx); call key_addi(l_maxx); call key_addi(l_miny); call key_addi(l_maxy); call key_addi(l_nk)
      call key_addi(F_i0); call key_addi(F_in); call key_addi(F_j0); call key_addi(F_jn); call key_addi(F_k0)
      call key_addi(merge(1, 0, Adz_BC_min_max_L)); call key_addi(merge(1, 0, Schm_autobar_L))
      do n = 1, 9
         call key_add(int(transfer(Adz_BC_pexp_list(n), 0_c_int32_t), c_int64_t))
      end do
      do n = 1, F_ntr_bc
         call key_addp(tr(n)%p); call key_addp(tr(n)%m); call key_addp(tr(n)%lin); call key_addp(tr(n)%mn)
         call key_addp(tr(n)%mx); call key_addi(tr(n)%legacy); call key_addi(tr(n)%weight); call key_addi(tr(n)%pexp_n)
      end do
      h = cache_find(K_BC)
      if (.not. c_associated(h)) then
         call adz_check(adz_bc_create(tr, int(F_ntr_bc, c_int), dom, int(Adz_field_dtype, c_int), h), 'adz_bc_create')
         call cache_put(K_BC, h)
      end if
      if (lam_flux_L) then  ! Bermejo_Conde.F90:70-72,92: the Aranami flux masses of each tracer
         call adz_check(adz_bermejo_conde_lam(h, airm1, airm0, pkps, mass_mask(), flux_mem, Adz_bc_deficit, &
                                              c_null_ptr, Adz_stream), 'adz_bermejo_conde_lam')
      else
         call adz_check(adz_bermejo_conde(h, airm1, airm0, pkps, mass_mask(), Adz_bc_deficit, c_null_ptr, &
                                          Adz_stream), 'adz_bermejo_conde')
      end if
      deallocate (tr)
   end subroutine Bermejo_Conde

end module adz_gpu
