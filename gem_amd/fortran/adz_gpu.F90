! adz_gpu.F90 -- the reference's adz_* entry points with their Fortran signatures, as thin
! ISO_C_BINDING calls into libadz_hip.so on device-resident fields and positions.
!
! Drop-in map (reference file:line -> routine here):
!   adz_set.F90:16-311          -> adz_gpu_set (grid description in, module state out)
!   adz_inittraj.F90:16-42      -> adz_inittraj
!   adz_traject.F90:16-130      -> adz_traject(F_dt_8)         (same signature)
!   adz_interp_hlt.F90:28-187   -> adz_tricub_hlt(F_stk, F_nptr, F_xyz, F_geom, F_num,
!                                    F_i0, F_in, F_j0, F_jn, F_k0, F_ext_L, F_post, F_QV_L)
!   adz_post_tr.F90:18-276      -> adz_post_tr(F_wp)            (same signature)
!   Bermejo_Conde.F90:18-320    -> Bermejo_Conde(F_bc, F_ntr_bc, F_i0, F_in, F_j0, F_jn, F_k0)
!   adz_BC_LAM_Aranami_hlt.F90:18-236 -> adz_BC_LAM_Aranami_hlt(F_stk, F_post, F_nptr), called by
!                                    adz_tricub_hlt with F_post on a LAM with Adz_BC_LAM_flux = 1
!   adz_BC_LAM_zlf_hlt.F90:18-319, adz_BC_LAM_zlf_0_hlt.F90:18-102 -> adz_BC_LAM_zlf_hlt(F1, F2, F_n,
!                                    F_setup), adz_BC_LAM_zlf_0_hlt(F_nptr, F_number)
!   ILMC_LAM.F90:18-273         -> ILMC_LAM(F_ns, F_name_S, F_i0, F_in, F_j0, F_jn), called by
!                                    adz_post_tr(1) for the tracers whose mono code is 2
!
! The module variables stand in for what the reference keeps in adz_mem / adz_options /
! dyn_fisl_options / cstv / tr3d (same names).  Array-valued members that the reference
! declares as real pointers (Adz_pntr_stack%src/dst, post%lin/min/max, bc%p/m/...) are
! type(c_ptr) device addresses of element (1,1,1) here, and the position sets Adz_pm,
! Adz_pmu, Adz_pmv, Adz_pt, Adz_pb are the device arrays the context owns.
!
! Differences a caller sees:
!   * F_xyz is the device address of the positions (type(c_ptr)), not a host real array.
!   * With F_post present, tracers whose meta has mono == 1 are clipped to [min, max]
!     inside the interpolation launch (the same operation adz_post_tr.F90:81-94 applies to
!     the same box afterwards), so adz_post_tr(1) only registers the fixer's tracers.
!   * The fixer of the BICHQV group (adz_post_tr(2)) is not part of this library: it stops
!     with a message.  ILMC (mono = 2), quintic vertical interpolation (F_QV_L, FP32
!     fields) and the ZLF boundary treatment (Adz_BC_LAM_flux = 2: adz_BC_LAM_zlf_0_hlt
!     around the interpolation on Adz_pb, the fixer on the Adz_pb box) are.
!   * Bermejo_Conde synchronises the device to return each tracer's mass_deficit in
!     F_bc(n)%mass_deficit, as the reference's host scalar.
!   * The LAM boundary flux keeps only what Bermejo_Conde reads of it: the masses of
!     FLUX_out / FLUX_in per tracer (mass_tr_PM.F90:101-127, with the air mass at time M
!     and geomh_mask_8 of the module state when adz_tricub_hlt runs), not the Adz_flux
!     arrays themselves.
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
   integer :: Schm_advec = 1
   real(kind=REAL64) :: Cstv_bA_m_8 = 0.6d0, Cstv_bA_8 = 0.6d0   ! GEM's default off-centring
   logical :: Adz_BC_min_max_L = .true.
   integer :: Adz_ILMC_sweep_max = 2             ! adz_options.F90:47-53
   logical :: Adz_ILMC_min_max_L = .true.
   logical :: Schm_autobar_L = .false.
   real :: Adz_BC_pexp_list(9) = 1.0
   integer :: Adz_pos_dtype = ADZ_F32       ! storage of positions: ADZ_F32 (the reference's real) / ADZ_F64
   integer :: Adz_field_dtype = ADZ_F32     ! storage of fields (the reference's real)
   integer :: Adz_interp_math = ADZ_MATH_REF
   type(c_ptr) :: Adz_stream = c_null_ptr   ! hipStream_t (null stream by default)
   ! winds read by adz_traject (adz_mem: Adz_uu_ext, Adz_vv_ext, Adz_ww_ext; dynamics: ut0, vt0, zdt0)
   type(c_ptr) :: Adz_uu_ext = c_null_ptr, Adz_vv_ext = c_null_ptr, Adz_ww_ext = c_null_ptr, &
                  ut0 = c_null_ptr, vt0 = c_null_ptr, zdt0 = c_null_ptr
   ! air mass at P and M, pr_k/pr_s and geomh_mask_8 read by Bermejo_Conde
   type(c_ptr) :: airm1 = c_null_ptr, airm0 = c_null_ptr, pkps = c_null_ptr, geomh_mask_8 = c_null_ptr
   ! tracer registry used by adz_post_tr (tr3d / mem_tracers / adz_mem)
   type(meta_tracers), allocatable, target :: Tr_3CWP(:)
   type(Adz_pntr_stack), allocatable :: Adz_stack(:)
   type(post), allocatable :: Adz_post(:)
   type(bc), allocatable :: Adz_bc(:)
   integer, private, save :: n_bc_TOTAL = 0
   ! masses (mass_fo, mass_fi) of the Aranami flux per tracer of the last adz_tricub_hlt
   real(kind=REAL64), allocatable :: Adz_flux_mass(:, :)

contains

   subroutine adz_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc /= ADZ_OK) then
         write (error_unit, '(a,": ",a," (status ",i0,")")') where, adz_error_message(), rc
         error stop 1
      end if
   end subroutine adz_check

   ! adz_set for one tile: builds the context and fills the module state.
   ! F_model_layout = (l_minx, l_maxx, l_miny, l_maxy) of the model (destination) arrays.
   subroutine adz_gpu_set(F_grid, F_itraj, F_model_layout)
      type(adz_grid), intent(in) :: F_grid
      integer, intent(in) :: F_itraj, F_model_layout(4)
      integer(c_int) :: ints(24), rc
      real(c_double) :: dbls(9)
      integer(c_long) :: ne
      if (c_associated(Adz_ctx)) call adz_check(adz_ctx_free(Adz_ctx), 'adz_ctx_free')
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
   end subroutine adz_gpu_set

   subroutine adz_gpu_finalize()
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

   subroutine adz_tricub_hlt(F_stk, F_nptr, F_xyz, F_geom, F_num, F_i0, F_in, F_j0, F_jn, F_k0, F_ext_L, F_post, &
                             F_QV_L)
      logical, intent(IN), optional :: F_QV_L, F_ext_L
      integer, intent(in) :: F_nptr, F_num, F_i0, F_in, F_j0, F_jn, F_k0
      type(c_ptr), intent(in) :: F_xyz
      type(meta_tracers), dimension(F_nptr), intent(in), optional :: F_post
      type(Adz_pntr_stack), dimension(F_nptr), intent(inout) :: F_stk
      type(c_ptr), intent(in) :: F_geom
      type(adz_field), allocatable :: fl(:)
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
      if ((F_in - F_i0 + 1) <= 0 .or. (F_jn - F_j0 + 1) <= 0) return
      ! adz_interp_hlt.F90:91-96: the LAM boundary flux of the fixer's tracers
      if (.not. ext_L .and. present(F_post)) then
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
            fl(cnt)%src = F_stk(n)%src
            fl(cnt)%dst = F_stk(n)%dst
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
         call adz_check(c_adz_tricub_hlt(fl, int(cnt, c_int), F_xyz, int(Adz_pos_dtype, c_int), F_geom, &
                                         int(F_i0, c_int), int(F_in, c_int), int(F_j0, c_int), int(F_jn, c_int), &
                                         int(F_k0, c_int), int(l_nk, c_int), int(l_minx, c_int), &
                                         int(l_maxx, c_int), int(l_miny, c_int), int(l_maxy, c_int), &
                                         merge(1_c_int, 0_c_int, ext_L), int(post_flag, c_int), &
                                         int(Adz_field_dtype, c_int), int(Adz_interp_math, c_int), Adz_stream), &
                        'adz_tricub_hlt')
      end do
      deallocate (fl)
   end subroutine adz_tricub_hlt

   ! adz_BC_LAM_Aranami_hlt.F90:18-236 for the tracers of F_post that use the fixer (mass 1
   ! or 111..139): FLUX_out / FLUX_in at Adz_pb and their masses into Adz_flux_mass(:, n).
   subroutine adz_BC_LAM_Aranami_hlt(F_stk, F_post, F_nptr)
      integer, intent(in) :: F_nptr
      type(Adz_pntr_stack), dimension(F_nptr), intent(in) :: F_stk
      type(meta_tracers), dimension(F_nptr), intent(in) :: F_post
      type(adz_flux_tracer), allocatable :: tr(:)
      type(adz_flux_domain) :: dom
      integer, allocatable :: tr_num(:)
      real(kind=REAL64), allocatable, target :: hm(:)
      type(c_ptr) :: h, d_mass
      integer(c_size_t) :: bytes
      integer :: n, nf
      if (allocated(Adz_flux_mass)) deallocate (Adz_flux_mass)
      allocate (Adz_flux_mass(2, F_nptr), tr(F_nptr), tr_num(F_nptr))
      Adz_flux_mass = 0.d0
      nf = 0
      do n = 1, F_nptr
         if (.not. (F_post(n)%mass == 1 .or. (F_post(n)%mass >= 111 .and. F_post(n)%mass <= 139))) cycle
         nf = nf + 1
         tr_num(nf) = n
         tr(nf)%src = F_stk(n)%src
      end do
      if (nf == 0) return
      if (.not. (c_associated(airm0) .and. c_associated(geomh_mask_8))) then
         write (error_unit, *) 'adz_BC_LAM_Aranami_hlt: airm0 and geomh_mask_8 must be set before the tracers'
         error stop 1
      end if
      dom = adz_flux_domain(l_minx, l_maxx, l_miny, l_maxy, l_nk, l_ni, l_nj, Adz_i0b, Adz_inb, Adz_j0b, Adz_jnb, &
                            1 + pil_w, l_ni - pil_e, 1 + pil_s, l_nj - pil_n, Adz_k0t)
      call adz_check(adz_flux_create(tr, int(nf, c_int), dom, Adz_cpntr_t, int(Adz_field_dtype, c_int), &
                                     int(Adz_pos_dtype, c_int), h), 'adz_flux_create')
      bytes = int(16, c_size_t) * int(nf, c_size_t)
      if (hipMalloc(d_mass, bytes) /= 0) call adz_check(ADZ_ENOMEM, 'adz_BC_LAM_Aranami_hlt: hipMalloc')
      call adz_check(c_adz_bc_lam_flux(h, Adz_pb, airm0, geomh_mask_8, d_mass, Adz_stream), 'adz_bc_lam_flux')
      allocate (hm(2 * nf))
      if (hipMemcpy(c_loc(hm), d_mass, bytes, hipMemcpyDeviceToHost) /= 0) &
         call adz_check(ADZ_EHIP, 'adz_BC_LAM_Aranami_hlt: hipMemcpy')
      do n = 1, nf
         Adz_flux_mass(:, tr_num(n)) = hm(2 * n - 1:2 * n)
      end do
      if (hipFree(d_mass) /= 0) call adz_check(ADZ_EHIP, 'adz_BC_LAM_Aranami_hlt: hipFree')
      call adz_check(adz_flux_free(h), 'adz_flux_free')
      deallocate (tr, tr_num, hm)
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

   ! F_wp = 1: ILMC on the Tr_3CWP tracers whose mono code is 2 (adz_post_tr.F90:104), then
   ! register those that use the fixer (the clipping of mono = 1 happened in adz_tricub_hlt);
   ! F_wp = 0: core limits, settings from each tracer's mass code, Bermejo_Conde, and the
   ! carried deficits back into Tr_3CWP.
   subroutine adz_post_tr(F_wp)
      integer, intent(in) :: F_wp
      integer :: n, num, i0, in, j0, jn, k0
      logical :: bc_L
      if (F_wp == 2) then
         write (error_unit, *) 'adz_post_tr: the BICHQV tracer group is not part of this library'
         error stop 1
      end if
      if (.not. allocated(Tr_3CWP)) return
      if (F_wp /= 0) then
         if (allocated(Adz_bc)) deallocate (Adz_bc)
         allocate (Adz_bc(size(Tr_3CWP)))
         n_bc_TOTAL = 0
         do n = 1, size(Tr_3CWP)
            if (Tr_3CWP(n)%mono == 2) call ILMC_LAM(n, 'TR/'//trim(Tr_3CWP(n)%name)//':M', Adz_i0, Adz_in, &
                                                    Adz_j0, Adz_jn)
            bc_L = Tr_3CWP(n)%mass == 1 .or. (Tr_3CWP(n)%mass >= 111 .and. Tr_3CWP(n)%mass <= 139)
            if (.not. bc_L) cycle
            n_bc_TOTAL = n_bc_TOTAL + 1
            Adz_bc(n_bc_TOTAL)%n = n
            Adz_bc(n_bc_TOTAL)%wp = F_wp
         end do
         return
      end if
      if (n_bc_TOTAL == 0) return
      i0 = 1 + pil_w; in = l_ni - pil_e; j0 = 1 + pil_s; jn = l_nj - pil_n; k0 = Adz_k0t
      if (Adz_BC_LAM_flux == 2) then  ! ZLF: the fixer's scope is the Adz_pb box (adz_post_tr.F90:147-151)
         i0 = Adz_i0b; in = Adz_inb; j0 = Adz_j0b; jn = Adz_jnb; k0 = 1
      end if
      do num = 1, n_bc_TOTAL
         n = Adz_bc(num)%n
         Adz_bc(num)%p = Adz_stack(n)%src
         Adz_bc(num)%m = Adz_stack(n)%dst
         Adz_bc(num)%lin = Adz_post(n)%lin
         Adz_bc(num)%min = Adz_post(n)%min
         Adz_bc(num)%max = Adz_post(n)%max
         Adz_bc(num)%name_S = 'TR/'//trim(Tr_3CWP(n)%name)//':M'
         Adz_bc(num)%mass_deficit = Tr_3CWP(n)%BC_mass_deficit
         Adz_bc(num)%LEGACY_L = Tr_3CWP(n)%mass == 1
         Adz_bc(num)%weight = 1
         Adz_bc(num)%pexp_n = 0
         if (.not. Adz_bc(num)%LEGACY_L) then
            Adz_bc(num)%weight = (Tr_3CWP(n)%mass - 100) / 10
            Adz_bc(num)%pexp_n = Tr_3CWP(n)%mass - (100 + Adz_bc(num)%weight * 10)
         end if
      end do
      call Bermejo_Conde(Adz_bc, n_bc_TOTAL, i0, in, j0, jn, k0)
      do num = 1, n_bc_TOTAL
         Tr_3CWP(Adz_bc(num)%n)%BC_mass_deficit = Adz_bc(num)%mass_deficit
      end do
   end subroutine adz_post_tr

   ! ILMC_LAM.F90:18-273 on Adz_stack(F_ns)%dst with Adz_post(F_ns)%min / %max and the air
   ! mass at time M (airm0), over F_i0:F_in x F_j0:F_jn x Adz_k0t:l_nk of a tile that touches
   ! the four LAM edges.  F_name_S names the tracer in messages.
   subroutine ILMC_LAM(F_ns, F_name_S, F_i0, F_in, F_j0, F_jn)
      character(len=*), intent(in) :: F_name_S
      integer, intent(in) :: F_ns, F_i0, F_in, F_j0, F_jn
      type(adz_ilmc_tracer) :: tr(1)
      type(adz_ilmc_domain) :: dom
      type(c_ptr) :: h
      if (Adz_field_dtype /= ADZ_F32) then  ! the ILMC kernels read and write `real` arrays
         write (error_unit, *) 'ILMC_LAM: needs FP32 fields (Adz_field_dtype = ADZ_F32) for ', trim(F_name_S)
         error stop 1
      end if
      if (.not. c_associated(airm0)) then
         write (error_unit, *) 'ILMC_LAM: airm0 must be set before ', trim(F_name_S)
         error stop 1
      end if
      tr(1) = adz_ilmc_tracer(Adz_stack(F_ns)%dst, Adz_post(F_ns)%min, Adz_post(F_ns)%max)
      dom = adz_ilmc_domain(l_minx, l_maxx, l_miny, l_maxy, l_nk, F_i0, F_in, F_j0, F_jn, Adz_k0t, &
                            Adz_ILMC_sweep_max, merge(1, 0, Adz_ILMC_min_max_L), merge(1, 0, Schm_autobar_L))
      call adz_check(adz_ilmc_create(tr, 1_c_int, dom, h), 'ILMC_LAM ('//trim(F_name_S)//'): adz_ilmc_create')
      call adz_check(adz_ilmc_lam(h, airm0, c_null_ptr, Adz_stream), 'ILMC_LAM ('//trim(F_name_S)//')')
      call adz_check(adz_ilmc_free(h), 'ILMC_LAM: adz_ilmc_free')
   end subroutine ILMC_LAM

   subroutine Bermejo_Conde(F_bc, F_ntr_bc, F_i0, F_in, F_j0, F_jn, F_k0)
      integer, intent(in) :: F_i0, F_in, F_j0, F_jn, F_k0, F_ntr_bc
      type(bc), intent(inout), dimension(F_ntr_bc) :: F_bc
      type(adz_bc_tracer), allocatable :: tr(:)
      type(adz_bc_domain) :: dom
      type(c_ptr) :: h, d_def, d_flux
      real(kind=REAL64), allocatable, target :: def(:), flux(:, :)
      integer(c_size_t) :: bytes
      integer :: n
      logical :: lam_flux_L
      if (F_ntr_bc <= 0) return
      ! Bermejo_Conde.F90:70-72,92: the Aranami flux masses of each tracer (zero otherwise)
      lam_flux_L = Adz_BC_LAM_flux == 1 .and. allocated(Adz_flux_mass)
      allocate (tr(F_ntr_bc), def(F_ntr_bc), flux(2, F_ntr_bc))
      flux = 0.d0
      do n = 1, F_ntr_bc
         tr(n)%p = F_bc(n)%p
         tr(n)%m = F_bc(n)%m
         tr(n)%lin = F_bc(n)%lin
         tr(n)%mn = F_bc(n)%min
         tr(n)%mx = F_bc(n)%max
         tr(n)%legacy = merge(1, 0, F_bc(n)%LEGACY_L)
         tr(n)%weight = F_bc(n)%weight
         tr(n)%pexp_n = F_bc(n)%pexp_n
         def(n) = F_bc(n)%mass_deficit
         if (lam_flux_L) then
            if (F_bc(n)%n >= 1 .and. F_bc(n)%n <= size(Adz_flux_mass, 2)) flux(:, n) = Adz_flux_mass(:, F_bc(n)%n)
         end if
      end do
      dom = adz_bc_domain(l_minx, l_maxx, l_miny, l_maxy, l_nk, F_i0, F_in, F_j0, F_jn, F_k0, &
                          merge(1, 0, Adz_BC_min_max_L), Adz_BC_pexp_list)
      bytes = int(8, c_size_t) * int(F_ntr_bc, c_size_t)
      if (hipMalloc(d_def, bytes) /= 0) call adz_check(ADZ_ENOMEM, 'Bermejo_Conde: hipMalloc')
      if (hipMemcpy(d_def, c_loc(def), bytes, hipMemcpyHostToDevice) /= 0) &
         call adz_check(ADZ_EHIP, 'Bermejo_Conde: hipMemcpy')
      call adz_check(adz_bc_create(tr, int(F_ntr_bc, c_int), dom, int(Adz_field_dtype, c_int), h), 'adz_bc_create')
      if (lam_flux_L) then
         if (hipMalloc(d_flux, 2 * bytes) /= 0) call adz_check(ADZ_ENOMEM, 'Bermejo_Conde: hipMalloc')
         if (hipMemcpy(d_flux, c_loc(flux), 2 * bytes, hipMemcpyHostToDevice) /= 0) &
            call adz_check(ADZ_EHIP, 'Bermejo_Conde: hipMemcpy')
         call adz_check(adz_bermejo_conde_lam(h, airm1, airm0, pkps, geomh_mask_8, d_flux, d_def, c_null_ptr, &
                                              Adz_stream), 'adz_bermejo_conde_lam')
      else
         call adz_check(adz_bermejo_conde(h, airm1, airm0, pkps, geomh_mask_8, d_def, c_null_ptr, Adz_stream), &
                        'adz_bermejo_conde')
      end if
      if (hipMemcpy(c_loc(def), d_def, bytes, hipMemcpyDeviceToHost) /= 0) &
         call adz_check(ADZ_EHIP, 'Bermejo_Conde: hipMemcpy')
      call adz_check(adz_bc_free(h), 'adz_bc_free')
      if (hipFree(d_def) /= 0) call adz_check(ADZ_EHIP, 'Bermejo_Conde: hipFree')
      if (lam_flux_L) then
         if (hipFree(d_flux) /= 0) call adz_check(ADZ_EHIP, 'Bermejo_Conde: hipFree')
      end if
      do n = 1, F_ntr_bc
         F_bc(n)%mass_deficit = def(n)
      end do
      deallocate (tr, def, flux)
   end subroutine Bermejo_Conde

end module adz_gpu
