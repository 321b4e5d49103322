! adz_f_driver.F90 -- a Fortran host program that runs one semi-Lagrangian tracer step
! through the drop-in routines of module adz_gpu, the way gemdyn's tstpdyn / adz_main /
! adz_tracers sequence calls them:
!   adz_set -> adz_inittraj -> adz_traject(dt) -> adz_tricub_hlt(tracers at Adz_pt, with
!   Adz_post, with the LAM boundary flux) -> adz_post_tr(1) -> adz_post_tr(0) (Bermejo-Conde)
!
! Usage: adz_f_driver <input file> <output file> [qv|zlf|graph]
!        (qv: the tracers with quintic vertical interpolation, F_QV_L = .true.;
!         graph: the tricubic step run once, then from the same start again as a hipGraph
!         captured on a created stream (hipStreamBeginCapture .. hipGraphLaunch): the
!         output is the replay's;
!         zlf: the ZLF boundary treatment, Adz_BC_LAM_flux = 2, as adz_tracers.F90:88-99,
!         154-163 runs it: interpolation on Adz_pb, the piloting values of the
!         destinations saved before and restored after the fixer)
!        adz_f_driver sizes      (prints the byte sizes of the bind(C) types)
!        adz_f_driver refuse_f64 ilmc|qv   (FP64 fields given to ILMC / the quintic path:
!                                 must stop with a message before any device call)
! The input is a flat little-endian stream (written by the caller):
!   int32  ni nj nk halox haloy pil itraj nfld mh pos_dtype
!   real64 inv_hx inv_hy inv_rayt zmin zmax wmstar_nk idz_t_nkm1 dt
!   real64 zm(0:nk+1) zt(0:nk+1) zx(0:nk) lat(1-haloy:nj+haloy)
!   real32 uu_ext vv_ext ww_ext ut0 vt0 zdt0          each (nit, njt, nk), extended layout
!   real32 tracers_P(nfld)                            each (ni+2mh, nj+2mh, nk), model layout
!   real32 airm1 airm0 pkps                            each model layout
!   real64 geomh_mask_8 (ni+2mh, nj+2mh)
!   int32  mass(nfld) mono(nfld)
! The output stream holds: the Adz_pt positions, the nfld interpolated (and fixed)
! tracers, Adz_post lin/min/max of every tracer, and the carried deficits.
program adz_f_driver
   use, intrinsic :: iso_c_binding
   use, intrinsic :: iso_fortran_env, only: REAL64, error_unit
   use adz_hip, only: adz_grid, ADZ_F32, ADZ_F64, hipMalloc, hipMemcpy, hipFree, hipDeviceSynchronize, &
                      hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipStreamCreate, hipStreamDestroy, &
                      hipStreamBeginCapture, hipStreamEndCapture, hipGraphInstantiate, hipGraphLaunch, &
                      hipGraphExecDestroy, hipGraphDestroy, hipStreamSynchronize
   use adz_gpu
   implicit none
   character(len=512) :: fin, fout, mode
   integer(c_int) :: hdr(10)
   real(c_double) :: dbl(8)
   integer :: ni, nj, nk, hx, hy, pil, itraj, nfld, mh, pdt, nit, njt, n, u
   integer(c_size_t) :: n_ext, n_mod, n_mask, npt
   real(c_double), allocatable, target :: zm(:), zt(:), zx(:), lat(:), mask(:)
   real(c_float), allocatable, target :: buf(:)
   real(c_float), allocatable, target :: posf(:)
   real(c_double), allocatable, target :: posd(:)
   integer(c_int), allocatable :: mass(:), mono(:)
   type(adz_grid), target :: g
   type(c_ptr), allocatable :: d_src(:), d_dst(:), d_lin(:), d_min(:), d_max(:), d_pil(:)
   real(kind=REAL64), allocatable :: deficits(:)
   type(c_ptr) :: graph, gexec

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   call get_command_argument(3, mode)
   if (trim(fin) == 'sizes') then
      call print_sizes()
      stop
   end if
   if (trim(fin) == 'refuse_f64') then  ! FP64 fields with a kernel that reads `real`: must stop
      Adz_field_dtype = ADZ_F64
      if (trim(fout) == 'qv') then
         allocate (Adz_stack(1))
         call adz_tricub_hlt(Adz_stack, 1, c_null_ptr, c_null_ptr, 1, 1, 1, 1, 1, 1, F_ext_L=.false., F_QV_L=.true.)
      else
         allocate (Adz_stack(1), Adz_post_3CWP(1))
         Adz_post => Adz_post_3CWP
         call ILMC_LAM(1, 'TR/X:M', 1, 1, 1, 1)
      end if
      write (*, '(a)') 'not refused'
      stop
   end if
   open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
   read (u) hdr
   read (u) dbl
   ni = hdr(1); nj = hdr(2); nk = hdr(3); hx = hdr(4); hy = hdr(5); pil = hdr(6); itraj = hdr(7)
   nfld = hdr(8); mh = hdr(9); pdt = hdr(10)
   nit = ni + 2 * hx; njt = nj + 2 * hy
   allocate (zm(nk + 2), zt(nk + 2), zx(nk + 1), lat(njt))
   read (u) zm, zt, zx, lat

   g%ni = ni; g%nj = nj; g%nk = nk; g%i0g = 1; g%j0g = 1; g%halox = hx; g%haloy = hy
   g%west = 1; g%east = 1; g%south = 1; g%north = 1
   g%pil_w = pil; g%pil_e = pil; g%pil_s = pil; g%pil_n = pil; g%gbpil_t = 0; g%bc_lam_flux = 1
   if (trim(mode) == 'zlf') g%bc_lam_flux = 2
   g%niu = ni - 1; g%njv = nj - 1
   g%inv_hx = dbl(1); g%inv_hy = dbl(2); g%inv_rayt = dbl(3)
   g%zm = c_loc(zm); g%zt = c_loc(zt); g%zx = c_loc(zx); g%lat = c_loc(lat)
   g%zmin = dbl(4); g%zmax = dbl(5); g%wmstar_nk = dbl(6); g%idz_t_nkm1 = dbl(7)

   Adz_pos_dtype = pdt
   Adz_field_dtype = ADZ_F32
   call adz_gpu_set(g, itraj, [1 - mh, ni + mh, 1 - mh, nj + mh])

   n_ext = int(nit, c_size_t) * njt * nk
   n_mod = int(ni + 2 * mh, c_size_t) * (nj + 2 * mh) * nk
   n_mask = int(ni + 2 * mh, c_size_t) * (nj + 2 * mh)
   allocate (buf(max(n_ext, n_mod)))
   Adz_uu_ext = upload_f32(n_ext); Adz_vv_ext = upload_f32(n_ext); Adz_ww_ext = upload_f32(n_ext)
   ut0 = upload_f32(n_ext); vt0 = upload_f32(n_ext); zdt0 = upload_f32(n_ext)

   allocate (d_src(nfld), d_dst(nfld), d_lin(nfld), d_min(nfld), d_max(nfld), d_pil(nfld))
   allocate (Adz_stack(nfld), Adz_post_3CWP(nfld), Tr_3CWP(nfld), mass(nfld), mono(nfld), deficits(nfld))
   Adz_post => Adz_post_3CWP
   do n = 1, nfld
      d_src(n) = upload_f32(n_mod)
      d_dst(n) = zeros_f32(n_mod); d_lin(n) = zeros_f32(n_mod)
      d_min(n) = zeros_f32(n_mod); d_max(n) = zeros_f32(n_mod); d_pil(n) = zeros_f32(n_mod)
   end do
   airm1 = upload_f32(n_mod); airm0 = upload_f32(n_mod); pkps = upload_f32(n_mod)
   allocate (mask(n_mask))
   read (u) mask
   geomh_mask_8 = upload_bytes(c_loc(mask), 8 * n_mask)
   read (u) mass, mono
   close (u)

   ! pointers address element (1,1,1) of each model-layout array (mh halo)
   do n = 1, nfld
      Adz_stack(n)%src = offset(d_src(n))
      Adz_stack(n)%dst = offset(d_dst(n))
      Adz_post(n)%lin = offset(d_lin(n))
      Adz_post(n)%min = offset(d_min(n))
      Adz_post(n)%max = offset(d_max(n))
      Adz_stack(n)%pil = offset(d_pil(n))
      Tr_3CWP(n)%mass = mass(n)
      Tr_3CWP(n)%mono = mono(n)
   end do
   airm1 = offset(airm1); airm0 = offset(airm0); pkps = offset(pkps)
   geomh_mask_8 = c_ptr_add(geomh_mask_8, 8_c_size_t * int(mh + (ni + 2 * mh) * mh, c_size_t))

   if (trim(mode) == 'graph') then
      if (hipStreamCreate(Adz_stream) /= 0) error stop 'hipStreamCreate'
   end if
   call adz_inittraj()
   call one_step()
   if (trim(mode) == 'graph') then
      ! the same step again from the same start: the trajectory and the carried deficits
      ! restart, then the step is captured once and the graph launched
      if (hipStreamSynchronize(Adz_stream) /= 0) error stop 'hipStreamSynchronize'
      call adz_inittraj()
      do n = 1, nfld
         Tr_3CWP(n)%BC_mass_deficit = 0.d0
      end do
      call adz_bc_upload_deficits()
      if (hipStreamBeginCapture(Adz_stream, 0_c_int) /= 0) error stop 'hipStreamBeginCapture'
      call one_step()
      if (hipStreamEndCapture(Adz_stream, graph) /= 0) error stop 'hipStreamEndCapture'
      if (hipGraphInstantiate(gexec, graph, c_null_ptr, c_null_ptr, 0_c_size_t) /= 0) error stop 'hipGraphInstantiate'
      if (hipGraphLaunch(gexec, Adz_stream) /= 0) error stop 'hipGraphLaunch'
      if (hipStreamSynchronize(Adz_stream) /= 0) error stop 'hipStreamSynchronize'
      if (hipGraphExecDestroy(gexec) /= 0 .or. hipGraphDestroy(graph) /= 0) error stop 'hipGraphDestroy'
   end if
   call adz_bc_sync_deficits()
   if (hipDeviceSynchronize() /= 0) error stop 'hipDeviceSynchronize'

   open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
   npt = 3_c_size_t * int(Adz_num_t, c_size_t)
   if (pdt == ADZ_F64) then
      allocate (posd(npt))
      call download(c_loc(posd), Adz_pt, 8 * npt)
      write (u) posd
   else
      allocate (posf(npt))
      call download(c_loc(posf), Adz_pt, 4 * npt)
      write (u) posf
   end if
   do n = 1, nfld
      call download(c_loc(buf), d_dst(n), 4 * n_mod); write (u) buf(1:n_mod)
   end do
   do n = 1, nfld
      call download(c_loc(buf), d_lin(n), 4 * n_mod); write (u) buf(1:n_mod)
      call download(c_loc(buf), d_min(n), 4 * n_mod); write (u) buf(1:n_mod)
      call download(c_loc(buf), d_max(n), 4 * n_mod); write (u) buf(1:n_mod)
   end do
   do n = 1, nfld
      deficits(n) = Tr_3CWP(n)%BC_mass_deficit
   end do
   write (u) deficits
   close (u)
   call adz_gpu_finalize()
   if (trim(mode) == 'graph') then
      if (hipStreamDestroy(Adz_stream) /= 0) error stop 'hipStreamDestroy'
   end if
   write (*, '(a,i0,a)') 'adz_f_driver: ', nfld, ' tracers advected and fixed'

contains

   ! trajectory, interpolation of the tracers at their positions with Adz_post (and the
   ! boundary flux), adz_post_tr(1) (ILMC, registry) and adz_post_tr(0) (the fixer)
   subroutine one_step()
      call adz_traject(dbl(8))
      if (trim(mode) == 'zlf') then
         call adz_BC_LAM_zlf_0_hlt(nfld, 0)
         call adz_tricub_hlt(Adz_stack, nfld, Adz_pb, Adz_cpntr_t, Adz_num_b, Adz_i0b, Adz_inb, Adz_j0b, Adz_jnb, &
                             1, .false., Tr_3CWP)
         call adz_BC_LAM_zlf_0_hlt(nfld, 1)
      else
         call adz_tricub_hlt(Adz_stack, nfld, Adz_pt, Adz_cpntr_t, Adz_num_t, Adz_i0, Adz_in, Adz_j0, Adz_jn, &
                             Adz_k0t, .false., Tr_3CWP, F_QV_L=trim(mode) == 'qv')
      end if
      call adz_post_tr(1)
      call adz_post_tr(0)
      if (trim(mode) == 'zlf') call adz_BC_LAM_zlf_0_hlt(nfld, 2)
   end subroutine one_step

   subroutine print_sizes()
      use adz_hip, only: adz_field, adz_item, adz_tune, adz_bc_tracer, adz_bc_domain, adz_flux_tracer, &
                         adz_flux_domain, adz_airmass_fields
      type(adz_field) :: f
      type(adz_item) :: it
      type(adz_tune) :: t
      type(adz_bc_tracer) :: b
      type(adz_bc_domain) :: d
      type(adz_flux_tracer) :: ft
      type(adz_flux_domain) :: fd
      type(adz_airmass_fields) :: am
      write (*, '(a,9(1x,i0))') 'sizes', c_sizeof(g), c_sizeof(f), c_sizeof(it), c_sizeof(t), c_sizeof(b), &
         c_sizeof(d), c_sizeof(ft), c_sizeof(fd), c_sizeof(am)
   end subroutine print_sizes

   function c_ptr_add(p, bytes) result(q)
      type(c_ptr), intent(in) :: p
      integer(c_size_t), intent(in) :: bytes
      type(c_ptr) :: q
      integer(c_intptr_t) :: a
      a = transfer(p, a) + int(bytes, c_intptr_t)
      q = transfer(a, q)
   end function c_ptr_add

   function offset(p) result(q)  ! element (1,1,1) of a (1-mh:ni+mh, 1-mh:nj+mh, nk) real array
      type(c_ptr), intent(in) :: p
      type(c_ptr) :: q
      q = c_ptr_add(p, 4_c_size_t * int(mh + (ni + 2 * mh) * mh, c_size_t))
   end function offset

   function upload_bytes(h, bytes) result(d)
      type(c_ptr), intent(in) :: h
      integer(c_size_t), intent(in) :: bytes
      type(c_ptr) :: d
      if (hipMalloc(d, bytes) /= 0) error stop 'hipMalloc'
      if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) /= 0) error stop 'hipMemcpy'
   end function upload_bytes

   function upload_f32(nel) result(d)  ! reads nel reals from the input stream into a new device array
      integer(c_size_t), intent(in) :: nel
      type(c_ptr) :: d
      read (u) buf(1:nel)
      d = upload_bytes(c_loc(buf), 4 * nel)
   end function upload_f32

   function zeros_f32(nel) result(d)
      integer(c_size_t), intent(in) :: nel
      type(c_ptr) :: d
      buf(1:nel) = 0.0
      d = upload_bytes(c_loc(buf), 4 * nel)
   end function zeros_f32

   subroutine download(h, d, bytes)
      type(c_ptr), intent(in) :: h, d
      integer(c_size_t), intent(in) :: bytes
      if (hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) /= 0) error stop 'hipMemcpy'
   end subroutine download

end program adz_f_driver
